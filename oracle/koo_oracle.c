/*
 * koo_oracle.c -- FP64 CPU oracle for the LMM Monte Carlo pricing path.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py, and only as the checker.  The product library
 * never links or calls it.
 *
 * Each function restates, in double precision and plain C, the reference
 * routine cited above it (paths relative to kooderive-code-r256/).  How each
 * one is pinned (golden vectors, the compiled reference gold in oracle/_ref,
 * closed forms or invariants) is recorded beside its test in
 * tests/test_oracle_*.py; functions with no such check are marked
 * "parity unpinned" here.
 */
#include "koo_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>


/* ======================================================================
 * Stage (a)
 * ==================================================================== */

/* Bratley-Fox direction numbers, 32 per dimension, Q0.32, from the Joe-Kuo
 * (degree, a, m[]) rows passed in (degree[d], a[d], m[d*max_m ..]).  The rows
 * are the caller's: the tests pass the table read out of the reference's own
 * compiled gold/sobol_primitives.cpp.
 * Follows gold/sobol_gold.cpp:118-164 (initSobolDirectionVectors). */
int orc_sobol_directions(int n_dims, const uint32_t *degree, const uint32_t *a_bits,
                         const uint32_t *m, int max_m, uint32_t *dirs) {
  if (n_dims < 1) return -1;
  for (int d = 0; d < n_dims; ++d) {
    uint32_t *v = dirs + 32 * d;
    if (d == 0) {
      for (int i = 0; i < 32; ++i) v[i] = 1u << (31 - i);
      continue;
    }
    int deg = (int)degree[d];
    uint32_t a = a_bits[d];
    if (deg < 1 || deg > max_m) return -1;
    for (int i = 0; i < deg; ++i) v[i] = m[(size_t)d * max_m + i] << (31 - i);
    for (int i = deg; i < 32; ++i) {
      uint32_t x = v[i - deg] ^ (v[i - deg] >> deg);
      for (int j = 1; j < deg; ++j)
        if ((a >> (deg - 1 - j)) & 1u) x ^= v[i - j];
      v[i] = x;
    }
  }
  return 0;
}

/* Sequential Gray-code Sobol: x_0 = 0, x_i = x_{i-1} ^ v[ctz(i)], output
 * points offset+1 .. offset+n.  Follows gold/sobol_gold.cpp:166-204
 * (sobolCPUints / sobolCPUintsOffset), skipping ahead with the Gray code
 * g(n) = n ^ (n >> 1) instead of generating the discarded prefix. */
void orc_sobol_ints(int n_vectors, int n_dims, uint32_t path_offset, const uint32_t *dirs,
                    uint32_t *out) {
  for (int d = 0; d < n_dims; ++d) {
    const uint32_t *v = dirs + 32 * d;
    uint32_t n = path_offset + 1u;
    uint32_t g = n ^ (n >> 1);
    uint32_t x = 0;
    for (int k = 0; k < 32; ++k)
      if ((g >> k) & 1u) x ^= v[k];
    uint32_t *o = out + (size_t)d * n_vectors;
    for (int i = 0; i < n_vectors; ++i) {
      o[i] = x;
      n += 1u;
      x ^= v[__builtin_ctz(n)];
    }
  }
}

/* Digital shift of the first number_dim rows.  gold/scramble_gold.cpp:11-24. */
void orc_scramble(uint32_t *data, const uint32_t *scramblers, int number_dim, int paths) {
  for (int d = 0; d < number_dim; ++d)
    for (int p = 0; p < paths; ++p) data[(size_t)d * paths + p] ^= scramblers[d];
}

/* Shaw-Brickman rational approximation of the normal quantile (W. T. Shaw and
 * N. Brickman, "Differential equations for Monte Carlo recycling and a
 * GPU-optimized normal quantile", 2009); coefficients as published and as
 * used by inverseCumulatives.h:146-172 and gold/math/normals_gold.h:191-216. */
static const double SB_P[6] = {1.2533136835212087879, 1.9797154223229267471,
                               0.80002295072483916762, 0.087403248265958578062,
                               0.0020751409553756572917, 4.744820732427972462e-6};
static const double SB_Q[6] = {1.0, 2.0795584360534589311, 1.2499328117341603014,
                               0.23668431621373705623, 0.0120098270559197768,
                               0.00010590620919921025259};

/* The reference's FP32 formula with its own u = float(x) * 2^-32 mapping
 * (inverseCumulatives.h:146-172), in float arithmetic, edge cases included
 * (x = 0 and x >= 2^32-128 give a non-finite value in the reference). */
float orc_invnorm_sb_ref_f32(uint32_t x) {
  const float P1 = 1.2533136835212087879f, P2 = 1.9797154223229267471f,
              P3 = 0.80002295072483916762f, P4 = 0.087403248265958578062f,
              P5 = 0.0020751409553756572917f, P6 = 4.744820732427972462e-6f;
  const float Q1 = 1.0f, Q2 = 2.0795584360534589311f, Q3 = 1.2499328117341603014f,
              Q4 = 0.23668431621373705623f, Q5 = 0.0120098270559197768f,
              Q6 = 0.00010590620919921025259f;
  volatile float v = (float)x * 2.3283064E-10F;
  int sgn = (v >= 0.5f);
  sgn = sgn - !sgn;
  float z = -logf(1.0f - ((float)sgn * ((2.0f * v) - 1.0f)));
  float p = P1 + z * (P2 + z * (P3 + z * (P4 + (P5 + P6 * z) * z)));
  float q = Q1 + z * (Q2 + z * (Q3 + z * (Q4 + (Q5 + Q6 * z) * z)));
  return (float)sgn * z * p / q;
}

/* Lower-tail probability t in (0, 1/2] and sign for the quantile solve. */
static double norm_cdf_lower(double x) { return 0.5 * erfc(-x * M_SQRT1_2); }

/* Accurate quantile of a lower-tail probability t in (0, 1/2]: Shaw-Brickman
 * start, then two Halley steps on Phi(x) = t using libm erfc.  Returns x <= 0. */
static double quantile_lower(double t) {
  if (t >= 0.5) return 0.0;
  double z = -log(2.0 * t);
  double p = SB_P[0] + z * (SB_P[1] + z * (SB_P[2] + z * (SB_P[3] + z * (SB_P[4] + z * SB_P[5]))));
  double q = SB_Q[0] + z * (SB_Q[1] + z * (SB_Q[2] + z * (SB_Q[3] + z * (SB_Q[4] + z * SB_Q[5]))));
  double x = -z * p / q;
  for (int it = 0; it < 2; ++it) {
    double e = norm_cdf_lower(x) - t;
    double dens = exp(-0.5 * x * x) / sqrt(2.0 * M_PI);
    double u = e / dens;
    x = x - u / (1.0 + 0.5 * x * u);
  }
  return x;
}

double orc_invnorm(double u) {
  if (u <= 0.0 || u >= 1.0) return u <= 0.0 ? -INFINITY : INFINITY;
  if (u < 0.5) return quantile_lower(u);
  return -quantile_lower(1.0 - u);
}

/* The product's uniform: u(x) = x * 2^-32, with x = 0 mapped to 2^-33; the
 * upper tail is taken as 1 - u = (2^32 - x) * 2^-32 exactly. */
double orc_invnorm_u32(uint32_t x) {
  const double two_m32 = 2.3283064365386962890625e-10;
  if (x == 0u) return quantile_lower(0.5 * two_m32);
  if (x < 0x80000000u) return quantile_lower((double)x * two_m32);
  if (x == 0x80000000u) return 0.0;
  return -quantile_lower((double)(0x100000000ull - (uint64_t)x) * two_m32);
}

void orc_invnorm_u32_array(const uint32_t *in, double *out, int64_t n) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) out[i] = orc_invnorm_u32(in[i]);
}

/* ======================================================================
 * Stage (b)
 * ==================================================================== */

/* Diagonal allocation of Sobol dimensions to (factor, step) pairs.
 * gold/Bridge_allocator_gold.cpp:38-76 (bridge_allocate_diagonal). */
void orc_bridge_allocate_diagonal(int factors, int steps, int *indices) {
  int i0 = 0, j0 = 0, i = 0, j = 0, counter = 0;
  while (counter < factors * steps) {
    indices[i * steps + j] = counter++;
    if (i == 0 || j == steps - 1) {
      if (i0 < factors - 1) {
        i0 += 1;
        j0 = 0;
      } else {
        i0 = factors - 1;
        j0 += 1;
      }
      i = i0;
      j = j0;
    } else {
      i -= 1;
      j += 1;
    }
  }
}

/* Jaeckel bridge construction on the unit grid t_i = i+1.
 * brownian_bridge.cu:355-367 and 639-700 (constructor and initialize()). */
void orc_bridge_tables(int steps, int *bridge_index, int *left_index, int *right_index,
                       double *left_weight, double *right_weight, double *std_dev) {
  double *t = (double *)malloc(sizeof(double) * steps);
  int *map = (int *)calloc(steps, sizeof(int));
  for (int i = 0; i < steps; ++i) t[i] = (double)(i + 1);
  map[steps - 1] = 1;
  bridge_index[0] = steps - 1;
  left_index[0] = right_index[0] = 0;
  std_dev[0] = sqrt(t[steps - 1]);
  left_weight[0] = right_weight[0] = 0.0;
  for (int j = 0, i = 1; i < steps; ++i) {
    while (map[j]) ++j;
    int k = j;
    while (!map[k]) ++k;
    int l = j + ((k - 1 - j) >> 1);
    map[l] = i;
    bridge_index[i] = l;
    left_index[i] = j;
    right_index[i] = k;
    if (j != 0) {
      left_weight[i] = (t[k] - t[l]) / (t[k] - t[j - 1]);
      right_weight[i] = (t[l] - t[j - 1]) / (t[k] - t[j - 1]);
      std_dev[i] = sqrt(((t[l] - t[j - 1]) * (t[k] - t[l])) / (t[k] - t[j - 1]));
    } else {
      left_weight[i] = (t[k] - t[l]) / t[k];
      right_weight[i] = t[l] / t[k];
      std_dev[i] = sqrt(t[l] * (t[k] - t[l]) / t[k]);
    }
    j = k + 1;
    if (j >= steps) j = 0;
  }
  free(map);
  free(t);
}

/* Bridge then successive differences; output (S,F,P).
 * brownian_bridge.cu:401-456 (transformCPU). */
void orc_bridge_transform(int steps, int factors, int paths, const int *indices,
                          const double *in, double *out) {
  int *bi = (int *)malloc(sizeof(int) * steps * 3);
  double *w = (double *)malloc(sizeof(double) * steps * 3);
  orc_bridge_tables(steps, bi, bi + steps, bi + 2 * steps, w, w + steps, w + 2 * steps);
#pragma omp parallel
  {
    double *path = (double *)malloc(sizeof(double) * steps);
#pragma omp for schedule(static)
    for (int p = 0; p < paths; ++p) {
      for (int f = 0; f < factors; ++f) {
        const int *idx = indices + f * steps;
        path[steps - 1] = w[2 * steps] * in[(size_t)idx[0] * paths + p];
        for (int i = 1; i < steps; ++i) {
          int j = bi[steps + i], k = bi[2 * steps + i], l = bi[i];
          double zz = in[(size_t)idx[i] * paths + p];
          if (j != 0)
            path[l] = w[i] * path[j - 1] + w[steps + i] * path[k] + w[2 * steps + i] * zz;
          else
            path[l] = w[steps + i] * path[k] + w[2 * steps + i] * zz;
        }
        for (int i = steps - 1; i >= 1; --i) path[i] -= path[i - 1];
        for (int i = 0; i < steps; ++i) out[((size_t)i * factors + f) * paths + p] = path[i];
      }
    }
    free(path);
  }
  free(bi);
  free(w);
}

/* The legacy power-of-two bridge on [S][vectors], S = 2^log2, differenced
 * output.  gold/Bridge_gold.h:168-252 (BrownianBridge<D>): last point
 * 2^(n/2) z_0; one-sided points 2^(n-i-1) - 1 = right/2 + sqrt(2^(n-i-1)/2) z;
 * then per level i = 2..n the midpoints j 2^(n-i) - 1 (j odd >= 3) =
 * (left + right)/2 + sqrt(2^(n-i)/2) z; point k takes the variate whose
 * number is its place in that construction order. */
void orc_bridge_pow2(int log2, int vectors, const double *in, double *out) {
  const int S = 1 << log2;
  int *var = (int *)malloc(sizeof(int) * S);
  int c = 0;
  for (int i = 0; i <= log2; ++i) {
    const int m = 1 << (log2 - i);
    for (int j = 1; j <= (1 << i); j += 2) var[j * m - 1] = c++;
  }
#pragma omp parallel
  {
    double *w = (double *)malloc(sizeof(double) * S);
#pragma omp for schedule(static)
    for (int p = 0; p < vectors; ++p) {
#define ORC_Z(k) in[(size_t)var[k] * vectors + p]
      w[S - 1] = ORC_Z(S - 1) * pow(2.0, 0.5 * log2);
      for (int i = 0; i < log2; ++i) {
        const int k = (1 << (log2 - i - 1)) - 1, right = (1 << (log2 - i)) - 1;
        w[k] = 0.5 * w[right] + sqrt(0.5 * pow(2.0, log2 - i - 1.0)) * ORC_Z(k);
      }
      for (int i = 2; i <= log2; ++i) {
        const int m = 1 << (log2 - i);
        for (int j = 3; j < (1 << i); j += 2) {
          const int k = j * m - 1;
          w[k] = 0.5 * w[k + m] + 0.5 * w[k - m] + sqrt(m * 0.5) * ORC_Z(k);
        }
      }
#undef ORC_Z
      for (int k = S - 1; k >= 1; --k) w[k] -= w[k - 1];
      for (int k = 0; k < S; ++k) out[(size_t)k * vectors + p] = w[k];
    }
    free(w);
  }
  free(var);
}

/* Sobol dimension i of path p to (variate v(i), factor f(i)) of an
 * (S, P, F) array, the orderings of BrownianBridgeMultiDim
 * (gold/Bridge_gold.h:291-331): 0 row (v = i / F, f = i % F), 1 column
 * (v = i % S, f = i / S), 2 triangular (anti-diagonal walk).
 * Bridge_multi_gpu.cu:67-112. */
void orc_bridge_reorder(const double *in, double *out, int log2, int factors, int paths,
                        int ordering) {
  const int S = 1 << log2, N = S * factors;
  int crow = 0, ccol = 0;
  for (int i = 0; i < N; ++i) {
    int v, f;
    if (ordering == 1) {
      v = i % S;
      f = i / S;
    } else if (ordering == 0) {
      v = i / factors;
      f = i % factors;
    } else {
      v = crow;
      f = ccol;
      if (ccol > 0 && crow < S - 1) {
        --ccol;
        ++crow;
      } else {
        const int rpc = crow + ccol + 1;
        ccol = rpc < factors - 1 ? rpc : factors - 1;
        crow = rpc - ccol;
      }
    }
    for (int p = 0; p < paths; ++p)
      out[((size_t)v * paths + p) * factors + f] = in[(size_t)i * paths + p];
  }
}

/* Cancellable swap under the LS multi-regression strategy, the second pass
 * cashFlowGeneratorEE<Swap, earlyExerciseNull, LSAMultiExerciseStrategy>
 * (gold/cashFlowGeneration_earlyEx_product.h:26-86 and its strategy
 * :405-640; GPU cashFlowGen_earlyExProduct_gpu.h:28-172, 756-1049).  The
 * strategy tables are the reference's packed GPU arrays (see the product
 * header); at depth d the estimate is the Andersen shift plus the regression
 * of (x_i - shift)/divisor; before the last depth a net value below the
 * date's lower cut exercises and above its upper cut continues, otherwise
 * the next depth refines; at the last depth exercise iff estimate < 0
 * (the exercise value of a cancellation).  Flows: swap flows
 * (aux = [K, payReceive, acc1[S], acc2[S]]) until cancellation or the end;
 * ex_index[p] = steps processed. */
static double orc_ls_multi_value(int cross, const double *fd, const int *id, int E, int ex,
                                 const double *sv, int sv_max, size_t P, int p, double ev,
                                 int *exercise) {
  const int depth_max = id[0], maxv = id[1], row = id[2], nv = id[3 + ex];
  const double *shifts = fd + E + (size_t)ex * depth_max * maxv;
  const double *divs = shifts + (size_t)E * depth_max * maxv;
  const double *lower = fd + E + 2 * (size_t)E * depth_max * maxv + (size_t)ex * depth_max;
  const double *upper = lower + (size_t)E * depth_max;
  const double *w0 = fd + E + 2 * (size_t)E * depth_max * maxv + 2 * (size_t)E * depth_max +
                     (size_t)ex * depth_max * row;
  const double *vars = sv + (size_t)ex * P * sv_max + p;
  double x[16];
  for (int depth = 0;;) {
    const double *w = w0 + (size_t)depth * row;
    for (int i = 0; i < nv; ++i) x[i] = (vars[(size_t)i * P] - shifts[maxv * depth + i]) / divs[maxv * depth + i];
    double value = fd[ex] + *w++;
    if (!cross || nv == 1) {
      for (int i = 0; i < nv; ++i) {
        value += (*w++) * x[i];
        value += (*w++) * x[i] * x[i];
      }
    } else if (nv == 2) {
      value += w[0] * x[0] + w[1] * x[1] + w[2] * x[0] * x[0] + w[3] * x[0] * x[1] +
               w[4] * x[1] * x[1];
    } else {
      for (int i = 0; i < nv; ++i) value += (*w++) * x[i];
      for (int i = 0; i < nv; ++i) value += (*w++) * x[i] * x[(i + 1) % nv];
      for (int i = 0; i < nv; ++i) value += (*w++) * x[i] * x[i];
    }
    ++depth;
    if (depth == depth_max) {
      *exercise = value < ev;
      return value;
    }
    if (value - ev < lower[depth - 1]) {
      *exercise = 1;
      return lower[depth - 1];
    }
    if (value - ev > upper[depth - 1]) {
      *exercise = 0;
      return upper[depth - 1];
    }
  }
}

void orc_cashflows_ee(int cross, double *f1, double *f2, const double *aux, const double *fd,
                      const int *id, int paths, int steps, int n_ex, const unsigned char *is_ex,
                      const double *r1, const double *sv, int sv_max, double *est,
                      int *ex_index) {
  const size_t P = (size_t)paths;
  for (int p = 0; p < paths; ++p) {
    int i = 0, ex = 0, done = 0, index = 0;
    for (; i < steps && !done; ++i) {
      const size_t o = (size_t)i * P + p;
      double a = 0.0, b = 0.0;
      int exercise = 0;
      if (is_ex[i]) {
        est[o] = orc_ls_multi_value(cross, fd, id, n_ex, ex, sv, sv_max, P, p, 0.0, &exercise);
        if (exercise) done = 1;
        ++ex;
      }
      if (!exercise) {
        b = aux[2 + steps + index] * r1[o] * aux[1];
        a = -aux[0] * aux[2 + index] * aux[1];
        ++index;
        done = index == steps;
      }
      f1[o] = a;
      f2[o] = b;
    }
    ex_index[p] = i;
    for (; i < steps; ++i) f1[(size_t)i * P + p] = f2[(size_t)i * P + p] = 0.0;
  }
}

/* y(s,r,p) = sum_f A(s,r,f) z(s,p,f) + mu_fixed(s,r) for r >= alive(s).
 * correlate_drift_gpu.cu:142-220 (correlation_drift_kernel); gold twin
 * gold/correlate_drift_gold.h:22-61 (which fills every r). */
void orc_correlate_drift(const double *in, double *out, const double *A, const int *alive,
                         const double *drift_fixed, int a_offset_per_step, int factors,
                         int out_dims, int paths, int steps) {
  for (int s = 0; s < steps; ++s)
    for (int p = 0; p < paths; ++p)
      for (int r = alive[s]; r < out_dims; ++r) {
        double t = 0.0;
        for (int f = 0; f < factors; ++f)
          t += A[(size_t)s * a_offset_per_step + r * factors + f] *
               in[((size_t)s * paths + p) * factors + f];
        out[((size_t)s * out_dims + r) * paths + p] = t + drift_fixed[s * out_dims + r];
      }
}

/* Spot-measure drift mu_r = sum_f A_rf sum_{j=alive..r} g_j A_jf,
 * g_j = (f_j + d_j) tau_j / (1 + f_j tau_j).
 * gold/LMM_evolver_gold.h:178-203 (spotDriftComputer::getDrifts), started at
 * the alive index. */
void orc_spot_drifts(const double *A_step, const double *rates, const double *taus,
                     const double *displacements, int factors, int n_rates, int alive,
                     double *drifts) {
  double e[64];
  for (int f = 0; f < factors; ++f) e[f] = 0.0;
  for (int r = 0; r < alive; ++r) drifts[r] = 0.0;
  for (int r = alive; r < n_rates; ++r) {
    double g = (rates[r] + displacements[r]) * taus[r] / (1.0 + rates[r] * taus[r]);
    double mu = 0.0;
    for (int f = 0; f < factors; ++f) {
      e[f] += A_step[r * factors + f] * g;
      mu += A_step[r * factors + f] * e[f];
    }
    drifts[r] = mu;
  }
}

/* Predictor-corrector log-Euler over all steps with fused discount ratios.
 * LMM_evolver_all.cu:506-747 (LMM_evolver_all_steps_pc_kernel_devicelog_discounts):
 * step 0 uses the host initial drifts, later steps the drifts of the previous
 * step's rates; only alive rates (r >= alive(s)) are written. */
void orc_lmm_pc_evolve(const double *initial_rates, const double *initial_log_rates,
                       const double *taus, const double *variates, const double *A,
                       const double *initial_drifts, const double *fixed_drifts,
                       const double *displacements, const int *alive, int paths, int factors,
                       int steps, int rates, int transposed_variates, double *rates_out,
                       double *log_rates_out, double *discounts_out) {
  (void)initial_rates;
  const size_t SP = (size_t)paths;
#pragma omp parallel for schedule(static)
  for (int p = 0; p < paths; ++p) {
    double e[64], ep[64];
    for (int s = 0; s < steps; ++s) {
      const int a = alive[s];
      const double *As = A + (size_t)s * rates * factors;
      double zf[64];
      for (int f = 0; f < factors; ++f)
        zf[f] = transposed_variates ? variates[((size_t)s * factors + f) * SP + p]
                                    : variates[((size_t)s * SP + p) * factors + f];
      for (int f = 0; f < factors; ++f) e[f] = ep[f] = 0.0;
      double df = 1.0;
      discounts_out[((size_t)s * (rates + 1) + a) * SP + p] = 1.0;
      for (int r = a; r < rates; ++r) {
        const double *Ar = As + r * factors;
        const double tau = taus[r], d = displacements[r];
        double mu, logf0;
        if (s == 0) {
          mu = initial_drifts[r];
          logf0 = initial_log_rates[r];
        } else {
          double fr = rates_out[((size_t)(s - 1) * rates + r) * SP + p];
          double g = (fr + d) * tau / (1.0 + fr * tau);
          mu = 0.0;
          for (int f = 0; f < factors; ++f) {
            e[f] += g * Ar[f];
            mu += Ar[f] * e[f];
          }
          logf0 = log_rates_out[((size_t)(s - 1) * rates + r) * SP + p];
        }
        double dW = 0.0;
        for (int f = 0; f < factors; ++f) dW += Ar[f] * zf[f];
        double lf = logf0 + mu + dW + fixed_drifts[s * rates + r];
        double Xp = exp(lf);
        double gp = Xp * tau / (1.0 + (Xp - d) * tau);
        double mup = 0.0;
        for (int f = 0; f < factors; ++f) {
          ep[f] += gp * Ar[f];
          mup += Ar[f] * ep[f];
        }
        lf += 0.5 * (mup - mu);
        double rate = exp(lf) - d;
        log_rates_out[((size_t)s * rates + r) * SP + p] = lf;
        rates_out[((size_t)s * rates + r) * SP + p] = rate;
        df /= (1.0 + rate * tau);
        discounts_out[((size_t)s * (rates + 1) + r + 1) * SP + p] = df;
      }
    }
  }
}

/* The multi-kernel evolvers on correlated increments incr (S,N,P) = A z +
 * fixed drift: LMM_evolver_pc_gpu (LMM_evolver_gpu.cu:366-605, one launch
 * per step) and LMM_evolver_euler_gpu (:835-1010).  Step 0 takes the initial
 * log-rates and drifts for every rate; step s >= 1 rebuilds the drift over
 * r >= alive[s] from the previous step's rates, g = (f tau + d tau) /
 * (1 + f tau), e_f += g a_f, mu = sum_f a_f e_f (the gold computes the same
 * sums from rate 0 with zero loadings for dead rates,
 * gold/LMM_evolver_gold.h:235-389).  PC: predictor with the predicted rates'
 * drift, log-rate += (mu^ - mu)/2.  Euler: the log of the displaced rate is
 * capped at log 10 (LMM_evolver_gpu.cu:20). */
void orc_lmm_steps(const double *initial_log_rates, const double *taus, const double *incr,
                   const double *A, const double *initial_drifts, const double *displacements,
                   const int *alive, int paths, int factors, int steps, int rates, int euler,
                   double *rates_out, double *log_rates_out) {
  const size_t SP = (size_t)paths;
  const double ceiling = log(10.0); /* the gold's ceiling; the GPU code's is its float rounding */
#pragma omp parallel for schedule(static)
  for (int p = 0; p < paths; ++p) {
    double e[64], ep[64];
    for (int s = 0; s < steps; ++s) {
      const int a = s == 0 ? 0 : alive[s];
      const double *As = A + (size_t)s * rates * factors;
      for (int f = 0; f < factors; ++f) e[f] = ep[f] = 0.0;
      for (int r = a; r < rates; ++r) {
        const double *Ar = As + r * factors;
        const double tau = taus[r], d = displacements[r];
        double mu, lf;
        if (s == 0) {
          mu = initial_drifts[r];
          lf = initial_log_rates[r];
        } else {
          const double fr = rates_out[((size_t)(s - 1) * rates + r) * SP + p];
          const double g = (fr * tau + d * tau) / (1.0 + fr * tau);
          mu = 0.0;
          for (int f = 0; f < factors; ++f) {
            e[f] += g * Ar[f];
            mu += Ar[f] * e[f];
          }
          lf = log_rates_out[((size_t)(s - 1) * rates + r) * SP + p];
        }
        lf += mu + incr[((size_t)s * rates + r) * SP + p];
        if (euler) {
          if (lf > ceiling) lf = ceiling;
        } else {
          const double xp = exp(lf);
          const double gp = xp * tau / (1.0 + (xp - d) * tau);
          double mup = 0.0;
          for (int f = 0; f < factors; ++f) {
            ep[f] += gp * Ar[f];
            mup += Ar[f] * ep[f];
          }
          lf += 0.5 * (mup - mu);
        }
        log_rates_out[((size_t)s * rates + r) * SP + p] = lf;
        rates_out[((size_t)s * rates + r) * SP + p] = exp(lf) - d;
      }
    }
  }
}

/* P(s,alive) = 1; P(s,r+1) = P(s,r) / (1 + tau_r f(s,r)).
 * gold/LMM_evolver_gold.h:460-492 (discount_ratios_computation_gold). */
void orc_discount_ratios(const double *rates, const double *taus, const int *alive, int paths,
                         int steps, int n_rates, double *discounts) {
  for (int s = 0; s < steps; ++s)
    for (int p = 0; p < paths; ++p) {
      double df = 1.0;
      discounts[((size_t)s * (n_rates + 1) + alive[s]) * paths + p] = df;
      for (int r = alive[s]; r < n_rates; ++r) {
        df /= 1.0 + taus[r] * rates[((size_t)s * n_rates + r) * paths + p];
        discounts[((size_t)s * (n_rates + 1) + r + 1) * paths + p] = df;
      }
    }
}

/* A(s,r) = sum_{j>=r} tau_j P(s,j+1).
 * gold/LMM_evolver_gold.h:494-522 (coterminal_annuity_ratios_computation_gold). */
void orc_coterminal_annuities(const double *discounts, const double *taus, const int *alive,
                              int paths, int steps, int n_rates, double *annuities) {
  for (int p = 0; p < paths; ++p)
    for (int s = 0; s < steps; ++s) {
      double ann = 0.0;
      for (int r = n_rates - 1; r >= alive[s]; --r) {
        ann += taus[r] * discounts[((size_t)s * (n_rates + 1) + r + 1) * paths + p];
        annuities[((size_t)s * n_rates + r) * paths + p] = ann;
      }
    }
}

/* SR(s,r) = (P(s,r) - P(s,N)) / A(s,r).
 * gold/LMM_evolver_gold.h:524-552 (coterminal_swap_rates_computation_gold). */
void orc_coterminal_swap_rates(const double *discounts, const double *annuities, const int *alive,
                               int paths, int steps, int n_rates, double *swap_rates) {
  for (int p = 0; p < paths; ++p)
    for (int s = 0; s < steps; ++s)
      for (int r = alive[s]; r < n_rates; ++r) {
        double ann = annuities[((size_t)s * n_rates + r) * paths + p];
        double df2 = discounts[((size_t)s * (n_rates + 1) + n_rates) * paths + p];
        double df1 = discounts[((size_t)s * (n_rates + 1) + r) * paths + p];
        swap_rates[((size_t)s * n_rates + r) * paths + p] = (df1 - df2) / ann;
      }
}

/* Single swap rate from discount ratios of one step, written at step_out.
 * gold/LMM_evolver_gold.h:554-584 (swap_rate_computation_gold). */
void orc_swap_rate(const double *discounts, const double *taus, int start_index, int end_index,
                   int paths, int step_in, int step_out, int n_rates, double *swap_rate) {
  for (int p = 0; p < paths; ++p) {
    const double *D = discounts + (size_t)step_in * (n_rates + 1) * paths + p;
    double num = D[(size_t)start_index * paths] - D[(size_t)end_index * paths];
    double den = 0.0;
    for (int r = start_index; r < end_index; ++r) den += taus[r] * D[(size_t)(r + 1) * paths];
    swap_rate[(size_t)step_out * paths + p] = num / den;
  }
}

/* B(0) = 1; B(i) = B(i-1) P(i-1,i-1) / P(i-1,i).
 * gold/LMM_evolver_gold.h:586-604 (spot_measure_numeraires_computation_gold). */
void orc_spot_numeraires(const double *discounts, int paths, int n_rates, double *numeraires) {
  for (int p = 0; p < paths; ++p) {
    double b = 1.0;
    numeraires[p] = b;
    for (int r = 1; r < n_rates; ++r) {
      const double *D = discounts + (size_t)(r - 1) * (n_rates + 1) * paths + p;
      b *= D[(size_t)(r - 1) * paths] / D[(size_t)r * paths];
      numeraires[(size_t)r * paths + p] = b;
    }
  }
}

/* out(s,p) = f(s, indices[s], p).
 * gold/LMM_evolver_gold.h:642-668 (forward_rate_extraction_gold). */
void orc_forward_extraction(const double *forwards, const int *indices, int paths, int steps,
                            int n_rates, double *out) {
  for (int s = 0; s < steps; ++s)
    for (int p = 0; p < paths; ++p)
      out[(size_t)s * paths + p] = forwards[((size_t)s * n_rates + indices[s]) * paths + p];
}

/* ======================================================================
 * Stage (c)
 * ==================================================================== */

/* Swap: aux = [K, payReceive, accrual1[S], accrual2[S]];
 * flow1 = -K acc1 pr, flow2 = acc2 rate1 pr.
 * cashFlowGeneration_product_gpu.h:128-188 (class Swap) with the driver loop of
 * cashFlowGeneration_gpu.h:23-77; gold twin
 * gold/cashFlowGeneration_product_gold.h. */
void orc_swap_flows(double *flows1, double *flows2, const double *aux, int paths, int steps,
                    const double *rates1) {
  double K = aux[0], pr = aux[1];
  for (int p = 0; p < paths; ++p)
    for (int s = 0; s < steps; ++s) {
      flows1[(size_t)s * paths + p] = -K * aux[2 + s] * pr;
      flows2[(size_t)s * paths + p] = aux[2 + steps + s] * rates1[(size_t)s * paths + p] * pr;
    }
}

/* TARN: aux = [totalCoupon, K, m, accrual1[S], accrual2[S]].
 * cashFlowGeneration_product_gpu.h:19-108 (class TARN). */
void orc_tarn_flows(double *flows1, double *flows2, const double *aux, int paths, int steps,
                    const double *rates1, const double *rates2) {
  double total = aux[0], K = aux[1], m = aux[2];
  const double *acc1 = aux + 3, *acc2 = aux + 3 + steps;
  for (int p = 0; p < paths; ++p) {
    double paid = 0.0;
    int done = 0, index = 0;
    for (int s = 0; s < steps; ++s) {
      double fl1 = 0.0, fl2 = 0.0;
      if (!done) {
        fl2 = acc2[index] * rates2[(size_t)s * paths + p];
        if (index + 1 == steps) {
          fl1 = paid - total;
          done = 1;
        } else {
          double unc = (K - m * rates1[(size_t)s * paths + p]) * acc1[index];
          if (unc < 0.0) unc = 0.0;
          double nt = paid + unc;
          if (nt >= total) {
            done = 1;
            fl1 = paid - total;
          } else {
            fl1 = -unc;
            paid = nt;
            ++index;
          }
        }
      }
      flows1[(size_t)s * paths + p] = fl1;
      flows2[(size_t)s * paths + p] = fl2;
    }
  }
}

/* Caplet strip (no reference counterpart; same interface as Swap):
 * aux = [K, accrual[S]]; flow1 = accrual max(rate1 - K, 0), flow2 = 0. */
void orc_caplet_flows(double *flows1, double *flows2, const double *aux, int paths, int steps,
                      const double *rates1) {
  double K = aux[0];
  for (int p = 0; p < paths; ++p)
    for (int s = 0; s < steps; ++s) {
      double x = rates1[(size_t)s * paths + p] - K;
      flows1[(size_t)s * paths + p] = aux[1 + s] * (x > 0.0 ? x : 0.0);
      flows2[(size_t)s * paths + p] = 0.0;
    }
}

/* Payment times -> bracketing rate-time indices and log-linear weights.
 * gold/cashFlowDiscounting_gold.h:210-245 (generateCashFlowIndicesAndWeights). */
void orc_cash_flow_indices_weights(const double *rate_times, const double *payment_times,
                                   int n_times, int *first_index, int *second_index,
                                   double *thetas) {
  for (int i = 0; i < n_times; ++i) {
    double t = payment_times[i];
    if (t < rate_times[0]) {
      first_index[i] = 0;
      second_index[i] = 1;
    } else if (t > rate_times[n_times - 1]) {
      first_index[i] = n_times - 1;
      second_index[i] = n_times;
    } else {
      int j = 1;
      while (t > rate_times[j]) ++j;
      first_index[i] = j - 1;
      second_index[i] = j;
    }
    thetas[i] = (t - rate_times[first_index[i]]) /
                (rate_times[second_index[i]] - rate_times[first_index[i]]);
  }
}

/* df = (P1/P_s)^(1-theta) (P2/P_s)^theta / B_s, plus the per-path sum.
 * gold/cashFlowDiscounting_gold.h:20-95 (cashFlowDiscounting_gold). */
void orc_cash_flow_discounting(const int *first_index, const int *second_index,
                               const double *thetas, const double *discounts,
                               const double *flows, const double *numeraires, int paths,
                               int steps, double *discounted, double *summed) {
  for (int p = 0; p < paths; ++p) {
    double total = 0.0;
    for (int s = 0; s < steps; ++s) {
      const double *D = discounts + (size_t)s * (steps + 1) * paths + p;
      double b = D[(size_t)s * paths];
      double r1 = D[(size_t)first_index[s] * paths] / b;
      double r2 = D[(size_t)second_index[s] * paths] / b;
      double df = pow(r1, 1.0 - thetas[s]) * pow(r2, thetas[s]);
      double cf = flows[(size_t)s * paths + p] * df / numeraires[(size_t)s * paths + p];
      discounted[(size_t)s * paths + p] = cf;
      total += cf;
    }
    summed[p] = total;
  }
}

/* Discount each flow to its generation step with the curve of that step, then
 * to targetIndices[i] with the numeraire ratio.
 * cashFlowDiscounting_gpu.cu:369-440 (cashFlowDiscounting_partial_kernel),
 * which divides the interpolated ratio once more by P(i,i) (== 1 for the
 * evolver's discount ratios); gold twin gold/cashFlowDiscounting_gold.h:97-205. */
void orc_cash_flow_discounting_partial(int flows_per_path, const int *gen_indices,
                                       const int *first_index, const int *second_index,
                                       const double *thetas, const int *target_indices,
                                       const double *discounts, const double *flows,
                                       const double *numeraires, int paths, int steps,
                                       double *discounted) {
  for (int p = 0; p < paths; ++p)
    for (int k = 0; k < flows_per_path; ++k) {
      int i = gen_indices[k];
      const double *D = discounts + (size_t)i * (steps + 1) * paths + p;
      double now = D[(size_t)i * paths];
      double r1 = D[(size_t)first_index[k] * paths] / now;
      double r2 = D[(size_t)second_index[k] * paths] / now;
      double rdf = pow(r1, 1.0 - thetas[k]) * pow(r2, thetas[k]);
      double fl = flows[(size_t)k * paths + p] * rdf / now;
      fl *= numeraires[(size_t)target_indices[k] * paths + p] / numeraires[(size_t)i * paths + p];
      discounted[(size_t)k * paths + p] = fl;
    }
}

/* Neumaier-compensated sum (reference sums are thrust::reduce in FP32;
 * the oracle sums exactly enough that its own error is negligible). */
double orc_sum_compensated(const double *x, int64_t n) {
  double s = 0.0, c = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    double t = s + x[i];
    if (fabs(s) >= fabs(x[i]))
      c += (s - t) + x[i];
    else
      c += (x[i] - t) + s;
    s = t;
  }
  return s + c;
}

/* ======================================================================
 * Host set-up: correlation and pseudo-roots
 * ==================================================================== */

/* Cyclic Jacobi eigen-decomposition of a symmetric matrix; eigenvalues in
 * descending order; each eigenvector's largest-magnitude component made
 * positive.  Plays the role of gold/math/Schur_gold.cpp
 * (SymmetricSchurDecomposition). */
int orc_sym_eigen(int n, const double *a, double *evals, double *evecs) {
  double *m = (double *)malloc(sizeof(double) * n * n);
  double *v = (double *)malloc(sizeof(double) * n * n);
  memcpy(m, a, sizeof(double) * n * n);
  for (int i = 0; i < n * n; ++i) v[i] = 0.0;
  for (int i = 0; i < n; ++i) v[i * n + i] = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, tot = 0.0;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        tot += m[i * n + j] * m[i * n + j];
        if (i != j) off += m[i * n + j] * m[i * n + j];
      }
    if (off <= 1e-30 * (tot > 0 ? tot : 1.0)) break;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        double apq = m[p * n + q];
        if (fabs(apq) < 1e-300) continue;
        double app = m[p * n + p], aqq = m[q * n + q];
        double theta = (aqq - app) / (2.0 * apq);
        double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; ++k) {
          double mkp = m[k * n + p], mkq = m[k * n + q];
          m[k * n + p] = c * mkp - s * mkq;
          m[k * n + q] = s * mkp + c * mkq;
        }
        for (int k = 0; k < n; ++k) {
          double mpk = m[p * n + k], mqk = m[q * n + k];
          m[p * n + k] = c * mpk - s * mqk;
          m[q * n + k] = s * mpk + c * mqk;
        }
        for (int k = 0; k < n; ++k) {
          double vkp = v[k * n + p], vkq = v[k * n + q];
          v[k * n + p] = c * vkp - s * vkq;
          v[k * n + q] = s * vkp + c * vkq;
        }
      }
  }
  int *order = (int *)malloc(sizeof(int) * n);
  for (int i = 0; i < n; ++i) order[i] = i;
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j)
      if (m[order[j] * n + order[j]] > m[order[i] * n + order[i]]) {
        int t = order[i];
        order[i] = order[j];
        order[j] = t;
      }
  for (int c = 0; c < n; ++c) {
    int src = order[c];
    evals[c] = m[src * n + src];
    int imax = 0;
    for (int k = 1; k < n; ++k)
      if (fabs(v[k * n + src]) > fabs(v[imax * n + src]) + 1e-12) imax = k;
    double sg = v[imax * n + src] < 0 ? -1.0 : 1.0;
    for (int k = 0; k < n; ++k) evecs[k * n + c] = sg * v[k * n + src];
  }
  free(order);
  free(v);
  free(m);
  return 0;
}

/* rho_ij = L + (1-L) exp(-beta |t_i - t_j|).
 * gold/volstructs_gold.cpp:157-175 (GenerateExpLongCorrMatrix). */
void orc_exp_corr(const double *times, int n, double L, double beta, double *corr) {
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j)
      corr[i * n + j] = (i == j) ? 1.0 : L + (1.0 - L) * exp(-beta * fabs(times[i] - times[j]));
}

/* Top-F eigenpairs scaled by sqrt(lambda), rows renormalised to keep the
 * diagonal.  gold/math/pseudoSquareRoot.cpp:76-110 (pseudoSqrtSchur). */
int orc_pseudo_sqrt(int n, const double *cov, int factors, int normalise, double *out) {
  double *ev = (double *)malloc(sizeof(double) * n);
  double *vec = (double *)malloc(sizeof(double) * n * n);
  orc_sym_eigen(n, cov, ev, vec);
  for (int f = 0; f < factors; ++f) {
    double rl = ev[f] > 0.0 ? sqrt(ev[f]) : 0.0;
    for (int i = 0; i < n; ++i) out[i * factors + f] = vec[i * n + f] * rl;
  }
  if (normalise)
    for (int i = 0; i < n; ++i) {
      double nrm = 0.0;
      for (int f = 0; f < factors; ++f) nrm += out[i * factors + f] * out[i * factors + f];
      if (nrm > 0.0) {
        double adj = sqrt(cov[i * n + i] / nrm);
        for (int f = 0; f < factors; ++f) out[i * factors + f] *= adj;
      }
    }
  free(vec);
  free(ev);
  return 0;
}

/* Per-step PCA pseudo-roots of the flat-vol covariance over each evolution
 * step.  gold/volstructs_gold.cpp:177-258 (FlatVolPseudoRootsOfCovariances),
 * correlation on rate start times rate_times[0..N-1]. */
int orc_flat_vol_pseudo_roots(const double *rate_times, const double *evol_times,
                              const double *vols, int n_rates, int steps, int factors, double L,
                              double beta, double *A) {
  double *corr = (double *)malloc(sizeof(double) * n_rates * n_rates);
  double *cov = (double *)malloc(sizeof(double) * n_rates * n_rates);
  orc_exp_corr(rate_times, n_rates, L, beta, corr);
  double start = 0.0;
  for (int s = 0; s < steps; ++s) {
    for (int r = 0; r < n_rates; ++r) {
      double tr = fmin(evol_times[s], rate_times[r]) - fmin(start, rate_times[r]);
      for (int q = 0; q <= r; ++q) {
        double tq = fmin(evol_times[s], rate_times[q]) - fmin(start, rate_times[q]);
        double c = corr[r * n_rates + q] * vols[r] * vols[q] * fmin(tr, tq);
        cov[r * n_rates + q] = cov[q * n_rates + r] = c;
      }
    }
    start = evol_times[s];
    orc_pseudo_sqrt(n_rates, cov, factors, 1, A + (size_t)s * n_rates * factors);
  }
  free(cov);
  free(corr);
  return 0;
}

/* Factor model A(s,r,f) = sigma_r sqrt(dt_s) B(r,f) for alive r (r >= s),
 * zero otherwise, with B the renormalised rank-F PCA root of the correlation
 * on rate start times; evolution times are the rate start times.  No
 * reference counterpart: the vol-row scaling makes d A / d sigma_r = A/sigma_r
 * exact, which is how vegas are defined. */
int orc_factor_pseudo_roots(const double *rate_times, const double *vols, int n_rates,
                            int factors, double L, double beta, double *A, double *B) {
  double *corr = (double *)malloc(sizeof(double) * n_rates * n_rates);
  double *b = (double *)malloc(sizeof(double) * n_rates * factors);
  orc_exp_corr(rate_times, n_rates, L, beta, corr);
  orc_pseudo_sqrt(n_rates, corr, factors, 1, b);
  double prev = 0.0;
  for (int s = 0; s < n_rates; ++s) {
    double dt = rate_times[s] - prev;
    prev = rate_times[s];
    for (int r = 0; r < n_rates; ++r)
      for (int f = 0; f < factors; ++f)
        A[((size_t)s * n_rates + r) * factors + f] =
            r >= s ? vols[r] * sqrt(dt) * b[r * factors + f] : 0.0;
  }
  if (B) memcpy(B, b, sizeof(double) * n_rates * factors);
  free(b);
  free(corr);
  return 0;
}

/* ======================================================================
 * Fused caplet-portfolio pipeline (price and pathwise adjoint).
 *
 * No reference counterpart: the reference lists adjoint Greeks as an open
 * task (doc/todo.txt:10).  The forward step is the A7 recurrence
 * (LMM_evolver_all.cu:571-742) with the step-0 drift computed from the
 * initial rates exactly as the host does for the reference kernel
 * (kooexample/bermudanIRDPricer.cu:446).  Product: caplet i fixes f_i at
 * step i and pays tau_i max(f_i - K, 0) at T_{i+1}; it is valued at T_i as
 * tau_i max(f_i - K, 0) / (1 + tau_i f_i) and deflated by the spot numeraire
 * B_i = prod_{j<i} (1 + tau_j f_j(T_j)) with B_0 = 1, so prices are in units
 * of the numeraire at T_0.
 * ==================================================================== */

typedef struct {
  double g, mu, gp, mup, x, xp, lp;
} step_rate;

/* One forward PC step from start log-state L (rates r >= s), writing the
 * end-of-step log-state into Lout and, if rec != NULL, the per-rate
 * intermediates and the running sums e(r), e^(r) (F each). */
static void pc_step(const orc_caplet_model *m, int s, const double *L, const double *z_s,
                    double *Lout, step_rate *rec, double *erec, double *eprec) {
  const int N = m->n_rates, F = m->factors;
  const double *As = m->A + (size_t)s * N * F;
  double e[64], ep[64];
  for (int f = 0; f < F; ++f) e[f] = ep[f] = 0.0;
  for (int r = s; r < N; ++r) {
    const double *Ar = As + r * F;
    double tau = m->taus[r], d = m->disp[r];
    double x = exp(L[r]);
    double g = tau * x / (1.0 + tau * (x - d));
    double mu = 0.0, dW = 0.0, c = 0.0;
    for (int f = 0; f < F; ++f) {
      e[f] += g * Ar[f];
      mu += Ar[f] * e[f];
      dW += Ar[f] * z_s[f];
      c -= 0.5 * Ar[f] * Ar[f];
    }
    double lh = L[r] + mu + c + dW;
    double xp = exp(lh);
    double gp = tau * xp / (1.0 + tau * (xp - d));
    double mup = 0.0;
    for (int f = 0; f < F; ++f) {
      ep[f] += gp * Ar[f];
      mup += Ar[f] * ep[f];
    }
    Lout[r] = lh + 0.5 * (mup - mu);
    if (rec) {
      rec[r].g = g;
      rec[r].mu = mu;
      rec[r].gp = gp;
      rec[r].mup = mup;
      rec[r].x = x;
      rec[r].xp = xp;
      rec[r].lp = lh;
      for (int f = 0; f < F; ++f) {
        erec[r * F + f] = e[f];
        eprec[r * F + f] = ep[f];
      }
    }
  }
}

static double caplet_value(double f, double tau, double K) {
  double x = f - K;
  return x > 0.0 ? tau * x / (1.0 + tau * f) : 0.0;
}

static double caplet_value_deriv(double f, double tau, double K) {
  if (f <= K) return 0.0;
  double den = 1.0 + tau * f;
  return tau * (1.0 + tau * K) / (den * den);
}

static double caplet_path_price(const orc_caplet_model *m, const double *z, int p, double *L,
                                double *L2) {
  const int N = m->n_rates, F = m->factors, P = m->paths;
  for (int r = 0; r < N; ++r) L[r] = log(m->f0[r] + m->disp[r]);
  double B = 1.0, y = 0.0, zs[64];
  for (int s = 0; s < N; ++s) {
    for (int f = 0; f < F; ++f) zs[f] = z[((size_t)s * F + f) * P + p];
    pc_step(m, s, L, zs, L2, NULL, NULL, NULL);
    for (int r = s; r < N; ++r) L[r] = L2[r];
    double fs = exp(L[s]) - m->disp[s];
    y += caplet_value(fs, m->taus[s], m->strike) / B;
    B *= 1.0 + m->taus[s] * fs;
  }
  return y;
}

void orc_caplet_price_paths(const orc_caplet_model *m, const double *z, double *price_per_path) {
  const int N = m->n_rates;
#pragma omp parallel
  {
    double *L = (double *)malloc(sizeof(double) * N * 2);
#pragma omp for schedule(static)
    for (int p = 0; p < m->paths; ++p) price_per_path[p] = caplet_path_price(m, z, p, L, L + N);
    free(L);
  }
}

/* Reverse sweep of one path; adds d(price)/d f0 into delta and
 * sum_{s,f} Abar(s,r,f) A(s,r,f) into vw (vega numerator); returns price. */
static double caplet_path_adjoint(const orc_caplet_model *m, const double *z, int p,
                                  double *states /* (S+1)*N */, double *lam, double *delta,
                                  double *vw, step_rate *rec, double *erec, double *eprec) {
  const int N = m->n_rates, F = m->factors, P = m->paths;
  double *Bs = states + (size_t)(N + 1) * N; /* N numeraires */
  for (int r = 0; r < N; ++r) states[r] = log(m->f0[r] + m->disp[r]);
  double B = 1.0, y = 0.0, zs[64];
  for (int s = 0; s < N; ++s) {
    for (int f = 0; f < F; ++f) zs[f] = z[((size_t)s * F + f) * P + p];
    double *Ls = states + (size_t)s * N, *Ln = states + (size_t)(s + 1) * N;
    for (int r = 0; r < N; ++r) Ln[r] = Ls[r];
    pc_step(m, s, Ls, zs, Ln, NULL, NULL, NULL);
    Bs[s] = B;
    double fs = exp(Ln[s]) - m->disp[s];
    y += caplet_value(fs, m->taus[s], m->strike) / B;
    B *= 1.0 + m->taus[s] * fs;
  }
  for (int r = 0; r < N; ++r) lam[r] = 0.0;
  double tail = 0.0;
  for (int s = N - 1; s >= 0; --s) {
    for (int f = 0; f < F; ++f) zs[f] = z[((size_t)s * F + f) * P + p];
    const double *Ls = states + (size_t)s * N;
    double Lend[64];
    pc_step(m, s, Ls, zs, Lend, rec, erec, eprec);
    /* payoff and numeraire adjoint of the rate fixing at the end of step s */
    double xs = exp(Lend[s]), fs = xs - m->disp[s], tau = m->taus[s];
    double v = caplet_value(fs, tau, m->strike);
    double dy = caplet_value_deriv(fs, tau, m->strike) / Bs[s] - tau / (1.0 + tau * fs) * tail;
    lam[s] += dy * xs;
    tail += v / Bs[s];
    /* reverse of the PC step */
    const double *As = m->A + (size_t)s * N * F;
    double Eh[64], E[64];
    for (int f = 0; f < F; ++f) Eh[f] = E[f] = 0.0;
    for (int r = N - 1; r >= s; --r) {
      const double *Ar = As + r * F;
      double d = m->disp[r];
      double lp = lam[r];
      step_rate q = rec[r];
      double gph = 0.0;
      for (int f = 0; f < F; ++f) {
        Eh[f] += 0.5 * lp * Ar[f];
        gph += Ar[f] * Eh[f];
      }
      double fh = q.xp - d;
      double lh = lp + gph * q.gp * (1.0 - m->taus[r] * d) / (1.0 + m->taus[r] * fh);
      double mub = lh - 0.5 * lp;
      double gb = 0.0;
      for (int f = 0; f < F; ++f) {
        E[f] += mub * Ar[f];
        gb += Ar[f] * E[f];
      }
      double fr = q.x - d;
      double lnew = lh + gb * q.g * (1.0 - m->taus[r] * d) / (1.0 + m->taus[r] * fr);
      double acc = 0.0;
      for (int f = 0; f < F; ++f) {
        double abar = 0.5 * lp * eprec[r * F + f] + q.gp * Eh[f] + mub * erec[r * F + f] +
                      q.g * E[f] + lh * zs[f] - lh * Ar[f];
        acc += abar * Ar[f];
      }
      vw[r] += acc;
      lam[r] = lnew;
    }
  }
  for (int r = 0; r < N; ++r) delta[r] += lam[r] / (m->f0[r] + m->disp[r]);
  return y;
}

void orc_caplet_adjoint_paths(const orc_caplet_model *m, const double *z, double *price_per_path,
                              double *delta_sum, double *vega_sum) {
  const int N = m->n_rates, F = m->factors;
  for (int r = 0; r < N; ++r) delta_sum[r] = vega_sum[r] = 0.0;
#pragma omp parallel
  {
    double *states = (double *)malloc(sizeof(double) * ((size_t)(N + 1) * N + N));
    double *lam = (double *)malloc(sizeof(double) * N);
    double *dl = (double *)calloc(N, sizeof(double));
    double *vw = (double *)calloc(N, sizeof(double));
    step_rate *rec = (step_rate *)malloc(sizeof(step_rate) * N);
    double *erec = (double *)malloc(sizeof(double) * N * F * 2);
#pragma omp for schedule(static)
    for (int p = 0; p < m->paths; ++p)
      price_per_path[p] = caplet_path_adjoint(m, z, p, states, lam, dl, vw, rec, erec, erec + N * F);
#pragma omp critical
    {
      for (int r = 0; r < N; ++r) {
        delta_sum[r] += dl[r];
        vega_sum[r] += vw[r] / m->sigma[r];
      }
    }
    free(erec);
    free(rec);
    free(vw);
    free(dl);
    free(lam);
    free(states);
  }
}

/* Integers -> FP64 normals -> diagonal-allocation bridge -> price (and
 * adjoint Greeks): the whole path in FP64 on the same Sobol points.  Outputs
 * are means over paths. */
void orc_caplet_pipeline(const orc_caplet_model *m, const uint32_t *ints, int do_adjoint,
                         double *price, double *deltas, double *vegas) {
  const int N = m->n_rates, F = m->factors, P = m->paths;
  const size_t D = (size_t)N * F;
  double *u = (double *)malloc(sizeof(double) * D * P);
  double *z = (double *)malloc(sizeof(double) * D * P);
  double *pp = (double *)malloc(sizeof(double) * P);
  int *idx = (int *)malloc(sizeof(int) * D);
  orc_invnorm_u32_array(ints, u, (int64_t)D * P);
  orc_bridge_allocate_diagonal(F, N, idx);
  orc_bridge_transform(N, F, P, idx, u, z);
  if (do_adjoint) {
    orc_caplet_adjoint_paths(m, z, pp, deltas, vegas);
    for (int r = 0; r < N; ++r) {
      deltas[r] /= P;
      vegas[r] /= P;
    }
  } else {
    orc_caplet_price_paths(m, z, pp);
  }
  *price = orc_sum_compensated(pp, P) / P;
  free(idx);
  free(pp);
  free(z);
  free(u);
}

/* ======================================================================
 * Stage (d): Longstaff-Schwartz
 * ==================================================================== */

/* Number of quadratic basis functions (quadraticPolynomialCrossGenerator,
 * gold/LS_regression.h:166-240, or the no-cross variant 1 + 2v). */
int orc_basis_count(int v, int use_cross) {
  if (!use_cross) return 1 + 2 * v;
  if (v == 0) return 1;
  if (v == 1) return 3;
  if (v == 2) return 6;
  return 1 + 3 * v;
}

/* Basis values: cross variant [1, x_i..., (v==1: x^2) (v==2: x^2, xy, y^2)
 * (v>=3: x_i x_{i+1 mod v}..., x_i^2...)] (gold/LS_regression.h:198-240);
 * no-cross variant [1, x_0, x_0^2, x_1, x_1^2, ...] (LS_gpu.h:92-126). */
void orc_basis_functions(const double *x, int v, int use_cross, double *out) {
  int j = 0;
  out[j++] = 1.0;
  if (!use_cross) {
    for (int i = 0; i < v; ++i) {
      out[j++] = x[i];
      out[j++] = x[i] * x[i];
    }
    return;
  }
  for (int i = 0; i < v; ++i) out[j++] = x[i];
  if (v == 1) {
    out[j++] = x[0] * x[0];
    return;
  }
  if (v == 2) {
    out[j++] = x[0] * x[0];
    out[j++] = x[0] * x[1];
    out[j++] = x[1] * x[1];
    return;
  }
  for (int i = 0; i < v; ++i) out[j++] = x[i] * x[(i + 1) % v];
  for (int i = 0; i < v; ++i) out[j++] = x[i] * x[i];
}

/* Cholesky solve of a symmetric positive-definite n x n system; returns 0
 * on success, -1 if a pivot is not positive. */
int orc_cholesky_solve(int n, const double *A, const double *b, double *x) {
  double L[64 * 64];
  if (n > 64) return -1;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j <= i; ++j) {
      double s = A[i * n + j];
      for (int k = 0; k < j; ++k) s -= L[i * 64 + k] * L[j * 64 + k];
      if (i == j) {
        if (!(s > 0.0)) return -1;
        L[i * 64 + i] = sqrt(s);
      } else {
        L[i * 64 + j] = s / L[j * 64 + j];
      }
    }
  double y[64];
  for (int i = 0; i < n; ++i) {
    double s = b[i];
    for (int k = 0; k < i; ++k) s -= L[i * 64 + k] * y[k];
    y[i] = s / L[i * 64 + i];
  }
  for (int i = n - 1; i >= 0; --i) {
    double s = y[i];
    for (int k = i + 1; k < n; ++k) s -= L[k * 64 + i] * x[k];
    x[i] = s / L[i * 64 + i];
  }
  return 0;
}

/* Solve of a symmetric n x n system by its eigendecomposition with small
 * eigenvalues dropped: x = sum over |l_i| >= tol * max|l| of
 * (v_i . b / l_i) v_i.  For a symmetric matrix this is the truncated SVD
 * solve of SVDSquareMatrix::SVDSolve (gold/math/svd_gold.cpp:560-596: keep
 * singular value s_i when s_i / s_0 >= tolerance), since the singular values
 * are the |l_i| and the singular vectors the v_i up to sign.  The
 * eigendecomposition is cyclic Jacobi (Golub and Van Loan, Matrix
 * Computations, 4th ed., 8.5.2, with the stable rotation of 8.5.1) run until
 * the off-diagonal mass is below 1e-30 of the total.  Returns 0. */
int orc_sym_solve_truncated(int n, const double *A, const double *b, double tol, double *x) {
  double a[64 * 64], v[64 * 64];
  if (n > 64 || n < 1) return -1;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      a[i * 64 + j] = A[i * n + j];
      v[i * 64 + j] = i == j ? 1.0 : 0.0;
    }
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, tot = 0.0;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        const double q = a[i * 64 + j] * a[i * 64 + j];
        tot += q;
        if (i != j) off += q;
      }
    if (off <= 1e-30 * tot) break;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        const double apq = a[p * 64 + q];
        if (apq == 0.0) continue;
        const double tau = (a[q * 64 + q] - a[p * 64 + p]) / (2.0 * apq);
        const double t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
        const double c = 1.0 / sqrt(1.0 + t * t), sn = t * c;
        for (int k = 0; k < n; ++k) {  /* A <- A J (columns p, q) */
          const double akp = a[k * 64 + p], akq = a[k * 64 + q];
          a[k * 64 + p] = c * akp - sn * akq;
          a[k * 64 + q] = sn * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {  /* A <- J^T A (rows p, q) */
          const double apk = a[p * 64 + k], aqk = a[q * 64 + k];
          a[p * 64 + k] = c * apk - sn * aqk;
          a[q * 64 + k] = sn * apk + c * aqk;
        }
        for (int k = 0; k < n; ++k) {  /* V <- V J */
          const double vkp = v[k * 64 + p], vkq = v[k * 64 + q];
          v[k * 64 + p] = c * vkp - sn * vkq;
          v[k * 64 + q] = sn * vkp + c * vkq;
        }
      }
  }
  double lmax = 0.0;
  for (int i = 0; i < n; ++i) lmax = fmax(lmax, fabs(a[i * 64 + i]));
  for (int i = 0; i < n; ++i) x[i] = 0.0;
  for (int i = 0; i < n; ++i) {
    const double l = a[i * 64 + i];
    if (!(lmax > 0.0) || fabs(l) < tol * lmax) continue;
    double e = 0.0;
    for (int k = 0; k < n; ++k) e += v[k * 64 + i] * b[k];
    e /= l;
    for (int k = 0; k < n; ++k) x[k] += e * v[k * 64 + i];
  }
  return 0;
}

/* The truncated solve of orc_sym_solve_truncated (eigenvalues below tol
 * times the largest dropped), reached through Cholesky: when A is positive
 * definite and 1 / trace(A^-1) >= tol * trace(A) -- a lower bound of the
 * smallest eigenvalue against an upper bound of the largest -- nothing is
 * dropped and the truncated solve is the Cholesky solution, which is
 * returned; otherwise the Jacobi eigen-solve with the cut.  (The GPU's LS
 * passes decide the same way.)  Returns 0. */
int orc_sym_solve_cut(int n, const double *A, const double *b, double tol, double *x) {
  double L[64 * 64];
  if (n > 64 || n < 1) return -1;
  int ok = 1;
  double trA = 0.0;
  for (int i = 0; i < n && ok; ++i) {
    trA += A[i * n + i];
    for (int j = 0; j <= i; ++j) {
      double s = A[i * n + j];
      for (int k = 0; k < j; ++k) s -= L[i * 64 + k] * L[j * 64 + k];
      if (i == j) {
        if (!(s > 0.0)) {
          ok = 0;
          break;
        }
        L[i * 64 + i] = sqrt(s);
      } else {
        L[i * 64 + j] = s / L[j * 64 + j];
      }
    }
  }
  if (ok) {
    double trInv = 0.0, z[64];
    for (int c = 0; c < n; ++c)
      for (int i = c; i < n; ++i) {
        double s = i == c ? 1.0 : 0.0;
        for (int q = c; q < i; ++q) s -= L[i * 64 + q] * z[q];
        z[i] = s / L[i * 64 + i];
        trInv += z[i] * z[i];
      }
    if (trInv * tol * trA <= 1.0) return orc_cholesky_solve(n, A, b, x);
  }
  return orc_sym_solve_truncated(n, A, b, tol, x);
}

/* Backward induction with one regression per exercise date (depth 0 of
 * generateRegressionCoefficientsViaLSMulti_gold, gold/MultiLS_regression.cpp:
 * 23-430, without variable normalisation): at date e the regression matrix is
 * the mean of phi phi^T and the target the mean of phi (cash_e + Y) over the
 * selected paths (gold/LS_regression.h:431-515), solved like the gold's SVD
 * with a 1e-8 cut-off (orc_sym_solve_cut: the Cholesky solution when a trace
 * bound shows no eigenvalue is below the cut, which is the usual case and
 * equals the truncated solve there; the truncated eigen-solve otherwise); then
 * Y = ex_e where ex_e exceeds the estimated continuation, else Y += cash_e
 * (updateDeflatedNextStepValuesMulti_gold), and Y is deflated to date e-1 by
 * B_{e-1}/B_e (updateDeflationOfContinuationValues_gold).  Price = mean of
 * Y / B_0.  With itm_only the regression uses only paths with ex_e > 0 and
 * exercise requires ex_e > 0 (the Bermudan configuration's rule); a date
 * with no such path gets zero coefficients. */
double orc_ls_first_pass(int n_ex, const double *bv, const int *vars_each_step, int max_vars,
                         const double *cash, const double *ev, const double *num, int paths,
                         int use_cross, int itm_only, double *coeffs) {
  const int maxd = orc_basis_count(max_vars, use_cross);
  double *Y = (double *)calloc(paths, sizeof(double));
  double x[16], phi[64], M[64 * 64], t[64];
  for (int e = n_ex - 1; e >= 0; --e) {
    const int v = vars_each_step[e], nd = orc_basis_count(v, use_cross);
    for (int i = 0; i < nd * nd; ++i) M[i] = 0.0;
    for (int i = 0; i < nd; ++i) t[i] = 0.0;
    long cnt = 0;
    for (int p = 0; p < paths; ++p) {
      const double exv = ev[(size_t)e * paths + p];
      if (itm_only && !(exv > 0.0)) continue;
      ++cnt;
      for (int i = 0; i < v; ++i) x[i] = bv[((size_t)e * max_vars + i) * paths + p];
      orc_basis_functions(x, v, use_cross, phi);
      const double target = cash[(size_t)e * paths + p] + Y[p];
      for (int i = 0; i < nd; ++i) {
        for (int j = 0; j <= i; ++j) M[i * nd + j] += phi[i] * phi[j];
        t[i] += phi[i] * target;
      }
    }
    double *beta = coeffs + (size_t)e * maxd;
    for (int i = 0; i < maxd; ++i) beta[i] = 0.0;
    if (cnt > 0) {
      for (int i = 0; i < nd; ++i) {
        for (int j = 0; j <= i; ++j) M[i * nd + j] = M[j * nd + i] = M[i * nd + j] / cnt;
        t[i] /= cnt;
      }
      orc_sym_solve_cut(nd, M, t, 1e-8, beta);
    }
    for (int p = 0; p < paths; ++p) {
      const double exv = ev[(size_t)e * paths + p];
      for (int i = 0; i < v; ++i) x[i] = bv[((size_t)e * max_vars + i) * paths + p];
      orc_basis_functions(x, v, use_cross, phi);
      double cont = 0.0;
      for (int i = 0; i < nd; ++i) cont += phi[i] * beta[i];
      const int allowed = !itm_only || exv > 0.0;
      if (allowed && exv > cont)
        Y[p] = exv;
      else
        Y[p] += cash[(size_t)e * paths + p];
    }
    if (e > 0)
      for (int p = 0; p < paths; ++p)
        Y[p] *= num[(size_t)(e - 1) * paths + p] / num[(size_t)e * paths + p];
  }
  double s = 0.0;
  for (int p = 0; p < paths; ++p) s += Y[p] / num[p];
  free(Y);
  return s / paths;
}

/* RegressionSelector_gold_Fraction::Select (Regression_Selector_concrete_gold.cpp:
 * 13-183, tolerance 1e-6): a symmetric band [-c, c] on the net values of the
 * pre-selected paths holding between lower_frac and upper_frac of them,
 * found from c = sd_guess * sd by growing (c /= multiplier), shrinking
 * (c *= multiplier) and bisection; marks the paths inside, returns their
 * count and the cuts.  Each of the three loops stops after
 * ORC_SELECT_MAX_STEPS steps (the gold's loops are unbounded and never end
 * when the net values have no spread, e.g. a zero standard deviation keeps
 * c = 0 under the growth step); where the gold ends, it ends within far
 * fewer steps, so the cap changes nothing there. */
#define ORC_SELECT_MAX_STEPS 1000
/* the gold's SVD tolerance for the LS normal equations
 * (gold/MultiLS_regression.cpp:172) */
#define ORC_SVD_TOLERANCE 1e-8
static long orc_count_band(const double *x, const int *pre, int n, double lo, double hi) {
  long c = 0;
  for (int p = 0; p < n; ++p)
    if (x[p] >= lo && x[p] <= hi && pre[p]) ++c;
  return c;
}

long orc_select_fraction(const double *net, const int *pre, int *sel, int paths, long n_pre,
                         double lower_frac, double upper_frac, double sd_guess, double multiplier,
                         double *lower_cut, double *upper_cut) {
  const long lowerT = (long)(n_pre * lower_frac), upperT = (long)(n_pre * upper_frac);
  double cut;
  if (n_pre > lowerT && n_pre < upperT) {
    cut = 10000.0;
  } else {
    double sum = 0.0, sumsq = 0.0;
    for (int p = 0; p < paths; ++p)
      if (pre[p]) {
        sum += net[p];
        sumsq += net[p] * net[p];
      }
    const double mean = sum / n_pre, sd = sqrt(sumsq / n_pre - mean * mean);
    double guess = sd * sd_guess;
    long count = orc_count_band(net, pre, paths, -guess, guess);
    int it = 0;
    while (count < lowerT && it++ < ORC_SELECT_MAX_STEPS) {
      guess /= multiplier;
      count = orc_count_band(net, pre, paths, -guess, guess);
    }
    if (count < upperT) {
      cut = guess;
    } else {
      double low = guess * multiplier;
      count = orc_count_band(net, pre, paths, -low, low);
      it = 0;
      while (count > upperT && it++ < ORC_SELECT_MAX_STEPS) {
        low *= multiplier;
        count = orc_count_band(net, pre, paths, -low, low);
      }
      if (count > lowerT) {
        cut = low;
      } else {
        double high = guess;
        guess = 0.5 * (low + high);
        count = orc_count_band(net, pre, paths, -guess, guess);
        it = 0;
        while ((count < lowerT || count > upperT) && (high - low) > 1e-6 &&
               it++ < ORC_SELECT_MAX_STEPS) {
          if (count < lowerT)
            low = guess;
          else
            high = guess;
          guess = 0.5 * (low + high);
          count = orc_count_band(net, pre, paths, -guess, guess);
        }
        cut = guess;
      }
    }
  }
  *lower_cut = -cut;
  *upper_cut = cut;
  long c = 0;
  for (int p = 0; p < paths; ++p) {
    sel[p] = net[p] >= -cut && net[p] <= cut && pre[p];
    c += sel[p];
  }
  return c;
}

/* RegressionSelectorStandardDeviations::Select_gold (RegressionSelectorConcrete.cu:
 * 50-90; the GPU form :19-46): cut = n_sds * sd of the net values over the
 * pre-selected paths (sd = sqrt(E[x^2] - mean^2)), keeping the paths strictly
 * inside (-cut, cut).  The reference rounds the cut to FP32; here it stays
 * FP64 like every other quantity of the pass.  Returns the count. */
long orc_select_sds(const double *net, const int *pre, int *sel, int paths, long n_pre,
                    double n_sds, double *lower_cut, double *upper_cut) {
  double sum = 0.0, sumsq = 0.0;
  for (int p = 0; p < paths; ++p)
    if (pre[p]) {
      sum += net[p];
      sumsq += net[p] * net[p];
    }
  const double mean = sum / n_pre, sd = sqrt(sumsq / n_pre - mean * mean);
  const double cut = n_sds * sd;
  *lower_cut = -cut;
  *upper_cut = cut;
  long c = 0;
  for (int p = 0; p < paths; ++p) {
    sel[p] = pre[p] && -cut < net[p] && net[p] < cut;
    c += sel[p];
  }
  return c;
}

/* The multi-regression first pass, generateRegressionCoefficientsViaLSMulti_gold
 * (gold/MultiLS_regression.cpp:20-420; GPU LS_main_cula.h:565-838).  Per
 * exercise date (last first): regression of cash + deflated next value on the
 * quadratic basis of the variables -- standardised by their mean and standard
 * deviation over the regression paths (depth 0: every path) when normalise is
 * set, raw otherwise (means 0, sds 1) -- ; net continuation values;
 * the selector's band (selector 0: RegressionSelector_gold_Fraction; 1:
 * RegressionSelectorStandardDeviations with sd_guess standard deviations)
 * picks the paths of the next regression, up to max_regressions, while more than
 * lower_path_cutoff paths remain (cuts 0 otherwise).  Then each path walks
 * the depths while its net value lies strictly inside the previous depth's
 * band (updateDeflatedNextStepValuesMulti_gold, gold/MultiLS_regression.h:
 * 47-116) and exercises when the exercise value exceeds the estimate.
 * Outputs (E x max_regressions x maxd coefficients, E x max_regressions
 * cuts, E x max_regressions x max_vars means and sds) are zero where a depth
 * was not reached.  Returns the first-pass price. */
double orc_ls_multi_first_pass(int n_ex, const double *bv, const int *vars_each_step,
                               int max_vars, const double *cash, const double *ev,
                               const double *num, int paths, int use_cross, int normalise,
                               int max_regressions, int lower_path_cutoff, int selector,
                               double lower_frac, double upper_frac,
                               double sd_guess, double multiplier, double *coeffs, double *lower,
                               double *upper, double *means, double *sds) {
  const int maxd = orc_basis_count(max_vars, use_cross), R = max_regressions;
  double *Y = (double *)calloc(paths, sizeof(double));
  double *net = (double *)calloc(paths, sizeof(double));
  int *pre = (int *)malloc(sizeof(int) * paths), *sel = (int *)calloc(paths, sizeof(int));
  double x[16], phi[64], M[64 * 64], t[64];
  for (size_t i = 0; i < (size_t)n_ex * R * maxd; ++i) coeffs[i] = 0.0;
  for (size_t i = 0; i < (size_t)n_ex * R; ++i) lower[i] = upper[i] = 0.0;
  for (size_t i = 0; i < (size_t)n_ex * R * max_vars; ++i) means[i] = sds[i] = 0.0;
  for (int e = n_ex - 1; e >= 0; --e) {
    const int v = vars_each_step[e], nd = orc_basis_count(v, use_cross);
    for (int p = 0; p < paths; ++p) pre[p] = 1;
    long n_pre = paths;
    for (int depth = 0; depth < R; ++depth) {
      double *m = means + ((size_t)e * R + depth) * max_vars;
      double *sd = sds + ((size_t)e * R + depth) * max_vars;
      for (int i = 0; i < v; ++i) {
        if (!normalise) {
          m[i] = 0.0;
          sd[i] = 1.0;
          continue;
        }
        double a = 0.0, b = 0.0;
        for (int p = 0; p < paths; ++p)
          if (pre[p]) {
            const double xv = bv[((size_t)e * max_vars + i) * paths + p];
            a += xv;
            b += xv * xv;
          }
        a /= n_pre;
        b /= n_pre;
        m[i] = a;
        sd[i] = sqrt(b - a * a);
      }
      for (int i = 0; i < nd * nd; ++i) M[i] = 0.0;
      for (int i = 0; i < nd; ++i) t[i] = 0.0;
      for (int p = 0; p < paths; ++p) {
        if (!pre[p]) continue;
        for (int i = 0; i < v; ++i) x[i] = (bv[((size_t)e * max_vars + i) * paths + p] - m[i]) / sd[i];
        orc_basis_functions(x, v, use_cross, phi);
        const double target = cash[(size_t)e * paths + p] + Y[p];
        for (int i = 0; i < nd; ++i) {
          for (int j = 0; j <= i; ++j) M[i * nd + j] += phi[i] * phi[j];
          t[i] += target * phi[i];
        }
      }
      for (int i = 0; i < nd; ++i) {
        M[i * nd + i] /= n_pre;
        t[i] /= n_pre;
        for (int j = 0; j < i; ++j) M[j * nd + i] = M[i * nd + j] = M[i * nd + j] / n_pre;
      }
      double *beta = coeffs + ((size_t)e * R + depth) * maxd;
      orc_sym_solve_truncated(nd, M, t, ORC_SVD_TOLERANCE, beta);
      for (int p = 0; p < paths; ++p) {
        if (!pre[p]) continue;
        for (int i = 0; i < v; ++i) x[i] = (bv[((size_t)e * max_vars + i) * paths + p] - m[i]) / sd[i];
        orc_basis_functions(x, v, use_cross, phi);
        double est = 0.0;
        for (int i = 0; i < nd; ++i) est += phi[i] * beta[i];
        net[p] = est - ev[(size_t)e * paths + p];
      }
      double lo, hi;
      const long after =
          selector == 1 ? orc_select_sds(net, pre, sel, paths, n_pre, sd_guess, &lo, &hi)
                        : orc_select_fraction(net, pre, sel, paths, n_pre, lower_frac, upper_frac,
                                              sd_guess, multiplier, &lo, &hi);
      if (after > lower_path_cutoff) {
        lower[(size_t)e * R + depth] = lo;
        upper[(size_t)e * R + depth] = hi;
      }
      if (after <= lower_path_cutoff) break;
      for (int p = 0; p < paths; ++p) pre[p] = sel[p];
      n_pre = after;
    }
    for (int p = 0; p < paths; ++p) {
      const double exv = ev[(size_t)e * paths + p];
      double est = 0.0, nv = 0.0;
      for (int depth = 0; depth == 0 || (depth < R && nv > lower[(size_t)e * R + depth - 1] &&
                                                    nv < upper[(size_t)e * R + depth - 1]);
           ++depth) {
        const double *m = means + ((size_t)e * R + depth) * max_vars;
        const double *sd = sds + ((size_t)e * R + depth) * max_vars;
        for (int i = 0; i < v; ++i) x[i] = (bv[((size_t)e * max_vars + i) * paths + p] - m[i]) / sd[i];
        orc_basis_functions(x, v, use_cross, phi);
        est = 0.0;
        for (int i = 0; i < nd; ++i) est += phi[i] * coeffs[((size_t)e * R + depth) * maxd + i];
        nv = est - exv;
      }
      if (exv > est)
        Y[p] = exv;
      else
        Y[p] += cash[(size_t)e * paths + p];
    }
    if (e > 0)
      for (int p = 0; p < paths; ++p)
        Y[p] *= num[(size_t)(e - 1) * paths + p] / num[(size_t)e * paths + p];
  }
  double s = 0.0;
  for (int p = 0; p < paths; ++p) s += Y[p] / num[p];
  free(Y);
  free(net);
  free(pre);
  free(sel);
  return s / paths;
}

/* Pricing under a frozen strategy: walk the dates forward, collect
 * cash_e / B_e until the strategy exercises, then add ex_e / B_e and stop
 * (SecondPassPriceLSAMultiUsingAggregatedFlows_gold,
 * gold/MultiLS_regression.h:254-300, depth 0). */
double orc_ls_second_pass(int n_ex, const double *bv, const int *vars_each_step, int max_vars,
                          const double *cash, const double *ev, const double *num, int paths,
                          int use_cross, int itm_only, const double *coeffs) {
  const int maxd = orc_basis_count(max_vars, use_cross);
  double total = 0.0;
#pragma omp parallel for reduction(+ : total) schedule(static)
  for (int p = 0; p < paths; ++p) {
    double x[16], phi[64], acc = 0.0;
    for (int e = 0; e < n_ex; ++e) {
      const int v = vars_each_step[e], nd = orc_basis_count(v, use_cross);
      const double exv = ev[(size_t)e * paths + p], b = num[(size_t)e * paths + p];
      for (int i = 0; i < v; ++i) x[i] = bv[((size_t)e * max_vars + i) * paths + p];
      orc_basis_functions(x, v, use_cross, phi);
      double cont = 0.0;
      for (int i = 0; i < nd; ++i) cont += phi[i] * coeffs[(size_t)e * maxd + i];
      const int allowed = !itm_only || exv > 0.0;
      if (allowed && cont - exv < 0.0) {
        acc += exv / b;
        break;
      }
      acc += cash[(size_t)e * paths + p] / b;
    }
    total += acc;
  }
  return total / paths;
}

/* ======================================================================
 * Bermudan payer swaption (the Longstaff-Schwartz workload).
 *
 * Exercise date e is evolution step t_e (rates r >= t_e alive).  From the
 * rates after that step: discount ratios P_{t_e} = 1,
 * P_{j+1} = P_j / (1 + tau_j f_j) (LMM_evolver_gold.h:460-492), co-terminal
 * annuity A = sum_{j>=t_e} tau_j P_{j+1} and swap rate SR = (1 - P_N) / A
 * (LMM_evolver_gold.h:494-560), exercise value max((SR - K) A, 0)
 * (earlyExerciseBermudanPayerSwaptionGold,
 * gold/early_exercise_value_generator_gold.h:93-124), basis variables
 * (SR, f_{t_e}) and the spot numeraire B_{t_e} = prod_{j<t_e} (1 + tau_j f_j(T_j))
 * (LMM_evolver_gold.h:600-640).  The product pays no cash flow before
 * exercise.
 * ==================================================================== */
void orc_bermudan_paths(const orc_caplet_model *m, const double *z, int n_ex, const int *ex_steps,
                        double *bv /* (E,2,P) */, double *ev /* (E,P) */, double *num /* (E,P) */) {
  const int N = m->n_rates, F = m->factors, P = m->paths;
#pragma omp parallel
  {
    double *L = (double *)malloc(sizeof(double) * N * 2), *L2 = L + N;
    double zs[64];
#pragma omp for schedule(static)
    for (int p = 0; p < P; ++p) {
      for (int r = 0; r < N; ++r) L[r] = log(m->f0[r] + m->disp[r]);
      double B = 1.0;
      int e = 0;
      for (int s = 0; s < N && e < n_ex; ++s) {
        for (int f = 0; f < F; ++f) zs[f] = z[((size_t)s * F + f) * P + p];
        pc_step(m, s, L, zs, L2, NULL, NULL, NULL);
        for (int r = s; r < N; ++r) L[r] = L2[r];
        const double fs = exp(L[s]) - m->disp[s];
        if (s == ex_steps[e]) {
          double disc = 1.0, ann = 0.0;
          for (int j = s; j < N; ++j) {
            disc /= 1.0 + m->taus[j] * (exp(L[j]) - m->disp[j]);
            ann += m->taus[j] * disc;
          }
          const double sr = (1.0 - disc) / ann;
          const double v = (sr - m->strike) * ann;
          const size_t o = (size_t)e * P + p;
          ev[o] = v > 0.0 ? v : 0.0;
          num[o] = B;
          bv[((size_t)e * 2 + 0) * P + p] = sr;
          bv[((size_t)e * 2 + 1) * P + p] = fs;
          ++e;
        }
        B *= 1.0 + m->taus[s] * fs;
      }
    }
    free(L);
  }
}

/* ======================================================================
 * Multi-dimensional Black-Scholes (basket Asian workload).
 * ==================================================================== */

/* Lower Cholesky factor L (row-major, L L^T = C); -1 if C is not positive
 * definite. */
int orc_cholesky_lower(int n, const double *C, double *L) {
  for (int i = 0; i < n * n; ++i) L[i] = 0.0;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j <= i; ++j) {
      double s = C[i * n + j];
      for (int k = 0; k < j; ++k) s -= L[i * n + k] * L[j * n + k];
      if (i == j) {
        if (!(s > 0.0)) return -1;
        L[i * n + i] = sqrt(s);
      } else {
        L[i * n + j] = s / L[j * n + j];
      }
    }
  return 0;
}

/* correlation_kernel_texture_multistep (Correlate_gpu.cu:93-114, device part
 * :26-90): out(s,r,p) = sum_f A(s,r,f) in(s,p,f) for r >= alive(s); the gold
 * twin (gold/Correlation_gold.h:22-43) starts its path loop at 1. */
void orc_correlated_paths(const double *in /* (S,P,F) */, double *out /* (S,N,P) */,
                          const double *A, const int *alive, int a_offset_per_step, int factors,
                          int out_dims, int paths, int steps) {
  for (int s = 0; s < steps; ++s) {
    const double *As = A + (size_t)s * a_offset_per_step;
    for (int p = 0; p < paths; ++p) {
      const double *z = in + ((size_t)s * paths + p) * factors;
      for (int r = alive[s]; r < out_dims; ++r) {
        double t = 0.0;
        for (int f = 0; f < factors; ++f) t += As[r * factors + f] * z[f];
        out[((size_t)s * out_dims + r) * paths + p] = t;
      }
    }
  }
}

/* multi_dim_BS_path_generator_nonlog_gold (LINUX/gold/MultiD_Path_Gen_BS_gold.h:
 * 44-101): cumulative log increments plus fixed drifts (S,N) (NULL: none, as
 * in the GPU kernel multid_path_gen_BS_gpu.cu:14-64) from the initial log
 * values, then exp. */
void orc_multid_bs_paths(const double *log0, const double *incr /* (S,N,P) */,
                         const double *drifts, double *out, int paths, int rates, int steps) {
  for (int p = 0; p < paths; ++p)
    for (int r = 0; r < rates; ++r) {
      double x = log0[r];
      for (int s = 0; s < steps; ++s) {
        const size_t o = ((size_t)s * rates + r) * paths + p;
        x += incr[o] + (drifts ? drifts[s * rates + r] : 0.0);
        out[o] = exp(x);
      }
    }
}

/* basketWeightings_gold (gold/weighted_average_gold.h:19-51). */
void orc_basket_weightings(const double *in /* (S,N,P) */, double *out /* (S,P) */,
                           const double *w, int paths, int stocks, int steps) {
  for (int p = 0; p < paths; ++p)
    for (int s = 0; s < steps; ++s) {
      double t = 0.0;
      for (int i = 0; i < stocks; ++i) t += w[i] * in[((size_t)s * stocks + i) * paths + p];
      out[(size_t)s * paths + p] = t;
    }
}

/* Arithmetic-average basket Asian call, no reference counterpart.  Per path:
 * x_a(t_{s+1}) = x_a(t_s) + (r - sigma_a^2/2) dt + sigma_a sqrt(dt) (L z_s)_a,
 * Abar = (1/S) sum_{s=1..S} sum_a w_a exp(x_a(t_s)), value e^{-rT} (Abar-K)^+.
 * Pathwise adjoint of the value: with d_a = sum_s S_a(t_s) and
 * v_a = sum_s S_a(t_s) dx_a(t_s)/dsigma_a,
 * dx_a(t_s)/dsigma_a = (x_a(t_s) - x_a(0) - (r + sigma_a^2/2) t_s) / sigma_a,
 * delta_a = e^{-rT} 1{Abar>K} w_a d_a / (S S_a(0)) and
 * vega_a = e^{-rT} 1{Abar>K} w_a v_a / S (L held fixed). */
static double basket_path(const orc_basket_model *m, const double *z, int p, double *x,
                          double *d, double *v, double *dl, double *vg) {
  const int A = m->assets, S = m->steps, P = m->paths;
  const double sq = sqrt(m->dt);
  for (int a = 0; a < A; ++a) {
    x[a] = m->x0[a];
    d[a] = v[a] = 0.0;
  }
  for (int s = 0; s < S; ++s) {
    const double t = (s + 1) * m->dt;
    for (int a = 0; a < A; ++a) {
      double c = 0.0;
      for (int f = 0; f <= a; ++f) c += m->chol[a * A + f] * z[((size_t)s * A + f) * P + p];
      const double sg = m->vols[a];
      x[a] += (m->rate - 0.5 * sg * sg) * m->dt + sg * sq * c;
      const double st = exp(x[a]);
      d[a] += st;
      v[a] += st * (x[a] - m->x0[a] - (m->rate + 0.5 * sg * sg) * t) / sg;
    }
  }
  double avg = 0.0;
  for (int a = 0; a < A; ++a) avg += m->weights[a] * d[a];
  avg /= S;
  const double disc = exp(-m->rate * S * m->dt);
  const int itm = avg > m->strike;
  for (int a = 0; a < A; ++a) {
    dl[a] = itm ? disc * m->weights[a] * d[a] / (S * exp(m->x0[a])) : 0.0;
    vg[a] = itm ? disc * m->weights[a] * v[a] / S : 0.0;
  }
  return itm ? disc * (avg - m->strike) : 0.0;
}

void orc_basket_asian_paths(const orc_basket_model *m, const double *z, double *price_per_path,
                            double *delta_sum, double *vega_sum) {
  const int A = m->assets;
  for (int a = 0; a < A; ++a) delta_sum[a] = vega_sum[a] = 0.0;
#pragma omp parallel
  {
    double *buf = (double *)malloc(sizeof(double) * 5 * A);
    double *dsum = (double *)calloc(2 * A, sizeof(double));
#pragma omp for schedule(static)
    for (int p = 0; p < m->paths; ++p) {
      price_per_path[p] = basket_path(m, z, p, buf, buf + A, buf + 2 * A, buf + 3 * A, buf + 4 * A);
      for (int a = 0; a < A; ++a) {
        dsum[a] += buf[3 * A + a];
        dsum[A + a] += buf[4 * A + a];
      }
    }
#pragma omp critical
    for (int a = 0; a < A; ++a) {
      delta_sum[a] += dsum[a];
      vega_sum[a] += dsum[A + a];
    }
    free(buf);
    free(dsum);
  }
}

/* ======================================================================
 * Bermudan swaption under a frozen exercise strategy: price and pathwise
 * adjoint Greeks (no reference counterpart; the strategy is the LS one of
 * the first pass, frozen as in the reference's second pass,
 * cashFlowGen_earlyExProduct_gpu.h:28-172).
 *
 * The path exercises at the first exercise date e with ev_e > 0 (when
 * itm_only) and ev_e > cont_e, cont_e = beta_e . phi((x - shift) * scale),
 * phi the quadratic basis with cross term in x = (SR, f_t).  Its value is
 * V = ev_tau / B_tau, ev = 1 - P_N - K A at T_t (t the exercise step),
 * P_{j+1} = P_j / (1 + tau_j f_j), A = sum_j tau_j P_{j+1}.  With the decision
 * held fixed: dV/df_j(T_t) = g_j (P_N + K Q_j) / B_t, g_j = tau_j / (1 + tau_j f_j),
 * Q_j = sum_{m > j} tau_{m-1} P_m, and dV/df_s(T_s) = -V g_s for the fixings
 * s < t inside B_t; then the PC steps t..0 are reversed as for the caplets.
 * ==================================================================== */
#define ORC_GIVEN_EV_TOL 1e-7
static double bermudan_path_adjoint(const orc_caplet_model *m, const orc_frozen_strategy *fs,
                                    const double *z, int p, double *states, double *lam,
                                    double *delta, double *vw, step_rate *rec, double *erec,
                                    double *eprec, int do_adjoint, int given, int *tau_out) {
  const int N = m->n_rates, F = m->factors, P = m->paths;
  double *Bs = states + (size_t)(N + 1) * N;
  for (int r = 0; r < N; ++r) states[r] = log(m->f0[r] + m->disp[r]);
  double B = 1.0, V = 0.0, zs[64];
  int t = -1, e = 0, kink_itm = 0;
  const int last = fs->ex_steps[fs->n_ex - 1];
  for (int s = 0; s <= last && t < 0; ++s) {
    for (int f = 0; f < F; ++f) zs[f] = z[((size_t)s * F + f) * P + p];
    double *Ls = states + (size_t)s * N, *Ln = states + (size_t)(s + 1) * N;
    for (int r = 0; r < N; ++r) Ln[r] = Ls[r];
    pc_step(m, s, Ls, zs, Ln, NULL, NULL, NULL);
    Bs[s] = B;
    const double fsx = exp(Ln[s]) - m->disp[s];
    if (s == fs->ex_steps[e]) {
      double disc = 1.0, ann = 0.0;
      for (int j = s; j < N; ++j) {
        disc /= 1.0 + m->taus[j] * (exp(Ln[j]) - m->disp[j]);
        ann += m->taus[j] * disc;
      }
      const double sr = (1.0 - disc) / ann;
      const double ev = (sr - m->strike) * ann;
      const double x[2] = {(sr - fs->shift[2 * e]) * fs->scale[2 * e],
                           (fsx - fs->shift[2 * e + 1]) * fs->scale[2 * e + 1]};
      double phi[6];
      orc_basis_functions(x, 2, 1, phi);
      double cont = 0.0;
      for (int i = 0; i < 6; ++i) cont += phi[i] * fs->coeffs[6 * e + i];
      const double evp = ev > 0.0 ? ev : 0.0;
      const int allowed = !fs->itm_only || evp > 0.0;
      /* given >= -1: the exercise step is the caller's (-1: never) */
      if (given >= -1 ? s == given : (allowed && cont - evp < 0.0)) {
        t = s;
        V = evp / B;
        /* A given exercise where the FP64 value is not positive but within
         * ORC_GIVEN_EV_TOL of zero -- below the FP32 resolution of
         * (SR - K) A, so an FP32 caller saw it in the money -- takes the
         * in-the-money side's derivative of the payoff's kink, as that
         * caller's adjoint does. */
        if (given >= -1 && !(ev > 0.0) && ev > -ORC_GIVEN_EV_TOL) kink_itm = 1;
      }
      ++e;
    }
    B *= 1.0 + m->taus[s] * fsx;
  }
  if (tau_out) *tau_out = t;
  if (!do_adjoint || t < 0 || !(V > 0.0 || kink_itm)) return V;
  /* payoff adjoint on the log-rates at the end of step t */
  for (int r = 0; r < N; ++r) lam[r] = 0.0;
  {
    const double *Le = states + (size_t)(t + 1) * N;
    double Pm[128];
    Pm[t] = 1.0;
    for (int j = t; j < N; ++j) Pm[j + 1] = Pm[j] / (1.0 + m->taus[j] * (exp(Le[j]) - m->disp[j]));
    double Q = 0.0;
    for (int j = N - 1; j >= t; --j) {
      const double fj = exp(Le[j]) - m->disp[j];
      const double g = m->taus[j] / (1.0 + m->taus[j] * fj);
      Q += m->taus[j] * Pm[j + 1]; /* Q_j = sum_{m > j} tau_{m-1} P_m */
      lam[j] = g * (Pm[N] + m->strike * Q) / Bs[t] * exp(Le[j]);
    }
  }
  for (int s = t; s >= 0; --s) {
    for (int f = 0; f < F; ++f) zs[f] = z[((size_t)s * F + f) * P + p];
    const double *Ls = states + (size_t)s * N;
    double Lend[64];
    pc_step(m, s, Ls, zs, Lend, rec, erec, eprec);
    if (s < t) { /* the fixing f_s(T_s) inside B_t */
      const double xs = exp(Lend[s]), fsx = xs - m->disp[s], tau = m->taus[s];
      lam[s] += -tau / (1.0 + tau * fsx) * V * xs;
    }
    const double *As = m->A + (size_t)s * N * F;
    double Eh[64], E[64];
    for (int f = 0; f < F; ++f) Eh[f] = E[f] = 0.0;
    for (int r = N - 1; r >= s; --r) {
      const double *Ar = As + r * F;
      const double d = m->disp[r];
      const double lp = lam[r];
      const step_rate q = rec[r];
      double gph = 0.0;
      for (int f = 0; f < F; ++f) {
        Eh[f] += 0.5 * lp * Ar[f];
        gph += Ar[f] * Eh[f];
      }
      const double fh = q.xp - d;
      const double lh = lp + gph * q.gp * (1.0 - m->taus[r] * d) / (1.0 + m->taus[r] * fh);
      const double mub = lh - 0.5 * lp;
      double gb = 0.0;
      for (int f = 0; f < F; ++f) {
        E[f] += mub * Ar[f];
        gb += Ar[f] * E[f];
      }
      const double fr = q.x - d;
      const double lnew = lh + gb * q.g * (1.0 - m->taus[r] * d) / (1.0 + m->taus[r] * fr);
      double acc = 0.0;
      for (int f = 0; f < F; ++f) {
        const double abar = 0.5 * lp * eprec[r * F + f] + q.gp * Eh[f] + mub * erec[r * F + f] +
                            q.g * E[f] + lh * zs[f] - lh * Ar[f];
        acc += abar * Ar[f];
      }
      vw[r] += acc;
      lam[r] = lnew;
    }
  }
  for (int r = 0; r < N; ++r) delta[r] += lam[r] / (m->f0[r] + m->disp[r]);
  return V;
}

static void bermudan_paths_adjoint(const orc_caplet_model *m, const orc_frozen_strategy *fs,
                                   const double *z, const int *given, int do_adjoint,
                                   double *value_per_path, int *exercise_step, double *delta_sum,
                                   double *vega_sum) {
  const int N = m->n_rates, F = m->factors;
  for (int r = 0; r < N; ++r) delta_sum[r] = vega_sum[r] = 0.0;
#pragma omp parallel
  {
    double *states = (double *)malloc(sizeof(double) * ((size_t)(N + 1) * N + N));
    double *lam = (double *)malloc(sizeof(double) * N);
    double *dl = (double *)calloc(N, sizeof(double)), *vw = (double *)calloc(N, sizeof(double));
    step_rate *rec = (step_rate *)malloc(sizeof(step_rate) * N);
    double *erec = (double *)malloc(sizeof(double) * N * F * 2), *eprec = erec + N * F;
#pragma omp for schedule(static)
    for (int p = 0; p < m->paths; ++p)
      value_per_path[p] = bermudan_path_adjoint(m, fs, z, p, states, lam, dl, vw, rec, erec, eprec,
                                                do_adjoint, given ? given[p] : -2,
                                                exercise_step ? exercise_step + p : NULL);
#pragma omp critical
    for (int r = 0; r < N; ++r) {
      delta_sum[r] += dl[r];
      vega_sum[r] += vw[r] / m->sigma[r];
    }
    free(states);
    free(lam);
    free(dl);
    free(vw);
    free(rec);
    free(erec);
  }
}

void orc_bermudan_frozen_paths(const orc_caplet_model *m, const orc_frozen_strategy *fs,
                               const double *z, int do_adjoint, double *value_per_path,
                               int *exercise_step, double *delta_sum, double *vega_sum) {
  bermudan_paths_adjoint(m, fs, z, NULL, do_adjoint, value_per_path, exercise_step, delta_sum,
                         vega_sum);
}

/* The same price and adjoint with each path's exercise step given
 * (given[p] = exercise step, -1 = never exercised): the GPU's own decisions
 * can be replayed in FP64, so its Greeks are compared on the same function
 * even where an FP32 continuation value sits within rounding of the
 * exercise value. */
void orc_bermudan_given_paths(const orc_caplet_model *m, const orc_frozen_strategy *fs,
                              const double *z, const int *given, int do_adjoint,
                              double *value_per_path, double *delta_sum, double *vega_sum) {
  bermudan_paths_adjoint(m, fs, z, given, do_adjoint, value_per_path, NULL, delta_sum, vega_sum);
}

/* Thread count of the OpenMP loops above (the CPU baseline times the oracle
 * on one core and on every core the process may run on). */
void orc_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
int orc_get_threads(void) { return omp_get_max_threads(); }
