// ls_multi.hip -- the multi-regression Longstaff-Schwartz first pass
// (generateMultiRegressionCoefficientsViaLS_gpu_flexisolve, LS_main_cula.h:
// 565-838, with RegressionSelectorFraction or RegressionSelectorStandardDeviations,
// RegressionSelectorConcrete.cu:19-233; the gold it is checked against is
// generateRegressionCoefficientsViaLSMultiQuadratic_gold,
// gold/MultiLS_regression.cpp:20-420).
//
// Per exercise date (last first) and regression depth d < R:
//   k_multi_prep    depth 0: the later date's exercise decision (walking its
//                   nested regressions) and the numeraire deflation of the
//                   running values Y; depth d > 0: membership of the paths
//                   selected by depth d-1's band.  FP64 partial sums of the
//                   basis variables over the members (normalisation).
//   k_multi_gram    means and standard deviations of the variables over the
//                   members; the Gram matrix of the normalised quadratic
//                   basis and the target (cash + Y) over the members on FP64
//                   MFMA (v_mfma_f64_16x16x4f64) into FP64 block partials.
//   k_multi_net     every block solves the normal equations (the gold's SVD
//                   solve with tolerance 1e-8: Cholesky when a trace bound
//                   shows no eigenvalue can fall under the tolerance, else a
//                   truncated Jacobi eigen-solve), then net continuation
//                   values of the members as 64-bit keys |net| (the FP64
//                   bits, which order like the values; non-members get a
//                   sentinel above every key) and FP64 sums of net and net^2.
//   k_multi_ladder  a histogram of the member keys against the ladder of cuts
//                   the selector's search can ask for (one pass, LDS bins).
//   k_multi_select  one block: the selector's band search (grow, shrink,
//                   bisect on a symmetric band) on the exact ladder counts;
//                   the cuts and the next depth's member count, or the
//                   interval of a bisection.
//   k_multi_tree    a histogram of the keys inside the bisection interval
//                   against every guess of the next ten bisection steps.
//   k_multi_bisect  one block: the bisection on those counts (two stages of
//                   tree + bisection cover twenty steps).
// The whole pass computes in FP64 from the FP32 inputs, so the band search
// takes the same steps as the FP64 gold: counts are exact and the net
// values agree to rounding.  Depths past the last active one return at once;
// the host launches every depth, so no result is read back during the pass.
#include <cstdlib>

#include "ls_common.h"

namespace koo {
namespace {

typedef double doublex4 __attribute__((ext_vector_type(4)));

constexpr int kMultiThreads = 1024;  // multi-block kernels (16 wavefronts)
constexpr int kSearchThreads = 1024;  // single-block search kernels (16 wavefronts)
constexpr int kMaxDepth = 8;         // regressions per exercise date
constexpr int kSelectMaxSteps = 1000;
constexpr unsigned long long kKeySentinel = 0x7FFFFFFFFFFFFFFFull;

__device__ __forceinline__ doublex4 mfma16x16x4_f64(double a, double b, doublex4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

struct MultiArgs {
  int e, E, P, max_vars, maxd, R, cross, normalise, cutoff;
  int selector;  // 0 fraction band, 1 sd_guess standard deviations
  int v, k;    // variables and basis functions at date e
  int vn, kn;  // at date e + 1 (when e + 1 < E)
  double lower_frac, upper_frac, sd_guess, multiplier;
};

// Values that one block writes and others read within a pass (partials,
// the search state, the outputs the later dates' decisions read, the keys)
// go through agent-coherent loads and stores, so the persistent pass needs
// no cache write-back or invalidation between its phases, only that each
// wave's stores are performed before its block counts itself in.
template <class T>
__device__ __forceinline__ T ld_ag(const T* p) {
  return __hip_atomic_load(const_cast<T*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ __forceinline__ void st_ag(T* p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Per depth of the current date: whether the depth is reached, its member
// count, and the key bound of the band of depth - 1 that defines its
// members (key <= cutbits).
struct DepthState {
  unsigned long long cutbits;
  long long n_pre;
  int active;
  int pad;
};

// Layout of the means and standard deviations kept in the workspace:
// norm[((e * R + d) * 2 + {0 mean, 1 sd}) * max_vars + i].
__device__ __forceinline__ size_t norm_at(int e, int R, int d, int which, int max_vars) {
  return ((static_cast<size_t>(e) * R + d) * 2 + which) * max_vars;
}

// Normalised variables of one path at (date q, depth d): (x - mean) / sd,
// FP64 as in the gold (GenerateLSRegressionMatrixSelectivelyNormalisingVariables_gold).
__device__ __forceinline__ void norm_vars(const float* __restrict__ bv, int q, int max_vars,
                                          size_t P, int p, int v, const double* m,
                                          const double* sd, double* x) {
#pragma unroll
  for (int i = 0; i < kMaxBasisVars; ++i)
    x[i] = i < v ? (static_cast<double>(bv[(static_cast<size_t>(q) * max_vars + i) * P + p]) -
                    m[i]) / sd[i]
                 : 0.0;
}

__device__ __forceinline__ double basis_dot(int v, int k, bool cross, const double* x,
                                            const double* beta) {
  double est = 0.0;
  for (int j = 0; j < k; ++j) est += basis_value(j, v, cross, x[0], x[1], x[2], x[3]) * beta[j];
  return est;
}

// The exercise decision of date q for one path (updateDeflatedNextStepValuesMulti_gold,
// gold/MultiLS_regression.h:47-116): walk the depths while the net value lies
// strictly inside the previous depth's band, exercise when the exercise
// value exceeds the last estimate, else add the date's cash flow.
// dec = the date's coefficients [R][maxd], norm [R][2][max_vars] and cuts
// [R] (lower, upper), in LDS.
__device__ __forceinline__ double decide(const MultiArgs& a, int q, int v, int k,
                                         const float* __restrict__ bv,
                                         const float* __restrict__ cash,
                                         const float* __restrict__ ev, size_t P, int p, double y,
                                         const double* beta, const double* nrm, const double* lo,
                                         const double* hi) {
  const double exv = static_cast<double>(ev[static_cast<size_t>(q) * P + p]);
  double est = 0.0, nv = 0.0;
  double x[kMaxBasisVars];
  for (int d = 0; d == 0 || (d < a.R && nv > lo[d - 1] && nv < hi[d - 1]); ++d) {
    norm_vars(bv, q, a.max_vars, P, p, v, nrm + (2 * d) * a.max_vars,
              nrm + (2 * d + 1) * a.max_vars, x);
    est = basis_dot(v, k, a.cross != 0, x, beta + d * a.maxd);
    nv = est - exv;
  }
  if (exv > est) return exv;
  return y + static_cast<double>(cash[static_cast<size_t>(q) * P + p]);
}

// Loads date q's decision data into LDS (every thread calls; contains a
// barrier).
__device__ __forceinline__ void load_decision(const MultiArgs& a, int q,
                                              const double* __restrict__ coeffs,
                                              const double* __restrict__ lower,
                                              const double* __restrict__ upper,
                                              const double* __restrict__ norm, double* beta,
                                              double* nrm, double* lo, double* hi) {
  for (int t = threadIdx.x; t < a.R * a.maxd; t += blockDim.x)
    beta[t] = ld_ag(coeffs + static_cast<size_t>(q) * a.R * a.maxd + t);
  for (int t = threadIdx.x; t < a.R * 2 * a.max_vars; t += blockDim.x)
    nrm[t] = ld_ag(norm + norm_at(q, a.R, 0, 0, a.max_vars) + t);
  for (int t = threadIdx.x; t < a.R; t += blockDim.x) {
    lo[t] = ld_ag(lower + static_cast<size_t>(q) * a.R + t);
    hi[t] = ld_ag(upper + static_cast<size_t>(q) * a.R + t);
  }
  __syncthreads();
}

// Block-wide FP64 sums of n values per lane (n <= 8) into
// part[t * nblocks + block].  Every lane calls it; contains barriers.
__device__ __forceinline__ void block_sums(const double* vals, int n, double* red /* [16][8] */,
                                           double* __restrict__ part, int block, int nblocks) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  for (int t = 0; t < n; ++t) {
    const double s = wave_sum_f64(vals[t]);
    if (lane == 0) red[wave * 8 + t] = s;
  }
  __syncthreads();
  if (threadIdx.x < n) {
    double s = 0.0;
    for (int w = 0; w < nw; ++w) s += red[w * 8 + threadIdx.x];
    st_ag(part + static_cast<size_t>(threadIdx.x) * nblocks + block, s);
  }
}

// ---------------------------------------------------------------------------
// Membership, decision and variable sums.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void multi_prep_body(const MultiArgs a, int depth, const float* __restrict__ bv,
                 const float* __restrict__ cash, const float* __restrict__ ev,
                 const float* __restrict__ num, const double* __restrict__ coeffs,
                 const double* __restrict__ lower, const double* __restrict__ upper,
                 const double* __restrict__ norm, double* __restrict__ Y,
                 unsigned char* __restrict__ mem, const unsigned long long* __restrict__ keys,
                 DepthState* __restrict__ state, double* __restrict__ stats_part, int blk, int nblk) {
  __shared__ double beta[kMaxDepth * kMaxGram];
  __shared__ double nrm[kMaxDepth * 2 * kMaxBasisVars];
  __shared__ double lo[kMaxDepth], hi[kMaxDepth];
  __shared__ double red[16 * 8];
  const size_t P = static_cast<size_t>(a.P);
  unsigned long long cutbits = 0;
  if (depth == 0) {
    if (blk == 0 && threadIdx.x == 0) {
      st_ag(&state[0].active, 1);
      st_ag(&state[0].n_pre, static_cast<long long>(a.P));
      st_ag(&state[0].cutbits, kKeySentinel);
    }
    if (a.e + 1 < a.E) load_decision(a, a.e + 1, coeffs, lower, upper, norm, beta, nrm, lo, hi);
  } else {
    if (!ld_ag(&state[depth].active)) return;  // uniform over the grid
    cutbits = ld_ag(&state[depth].cutbits);
  }
  double sums[2 * kMaxBasisVars];
#pragma unroll
  for (int i = 0; i < 2 * kMaxBasisVars; ++i) sums[i] = 0.0;
  for (int p = blk * blockDim.x + threadIdx.x; p < a.P; p += nblk * blockDim.x) {
    bool member;
    if (depth == 0) {
      double y = 0.0;
      if (a.e + 1 < a.E) {
        y = decide(a, a.e + 1, a.vn, a.kn, bv, cash, ev, P, p, Y[p], beta, nrm, lo, hi);
        // updateDeflationOfContinuationValues_gold (gold/LS_regression.cpp:98-115)
        y *= static_cast<double>(num[static_cast<size_t>(a.e) * P + p]) /
             static_cast<double>(num[static_cast<size_t>(a.e + 1) * P + p]);
      }
      Y[p] = y;
      member = true;
    } else {
      member = ld_ag(keys + p) <= cutbits;
    }
    mem[p] = member ? 1 : 0;
    if (a.normalise && member) {
#pragma unroll
      for (int i = 0; i < kMaxBasisVars; ++i)
        if (i < a.v) {
          const double x =
              static_cast<double>(bv[(static_cast<size_t>(a.e) * a.max_vars + i) * P + p]);
          sums[2 * i] += x;
          sums[2 * i + 1] += x * x;
        }
    }
  }
  if (a.normalise && a.v > 0)
    block_sums(sums, 2 * a.v, red, stats_part, blk, nblk);
}

__global__ void __launch_bounds__(kMultiThreads)
    k_multi_prep(const MultiArgs a, int depth, const float* __restrict__ bv,
                 const float* __restrict__ cash, const float* __restrict__ ev,
                 const float* __restrict__ num, const double* __restrict__ coeffs,
                 const double* __restrict__ lower, const double* __restrict__ upper,
                 const double* __restrict__ norm, double* __restrict__ Y,
                 unsigned char* __restrict__ mem, const unsigned long long* __restrict__ keys,
                 DepthState* __restrict__ state, double* __restrict__ stats_part) {
  multi_prep_body(a, depth, bv, cash, ev, num, coeffs, lower, upper, norm, Y, mem, keys, state, stats_part, blockIdx.x, gridDim.x);
}


// ---------------------------------------------------------------------------
// Means and standard deviations; FP64 MFMA Gram of basis and target.
// ---------------------------------------------------------------------------

// Adds the two diagonal 8x8 blocks (GROUPS == 2) or takes the 16x16 result
// of each wave's FP64 accumulator and sums the waves into the block's column
// of the entry-major partials part[(i*n + j) * nblocks + block].
// v_mfma_f64_16x16x4f64 takes A and B like the FP32 16x16x4 form (lane l
// supplies A[l%16][l/16] and B[l/16][l%16]) but interleaves the rows of D:
// element r of lane l is D[l/16 + 4r][l%16] (measured on gfx950).
template <int GROUPS>
__device__ __forceinline__ void block_gram_f64(const doublex4& acc, int n, double* lds_d,
                                               double* __restrict__ part, int block,
                                               int nblocks) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  double* d = lds_d + wave * 256;
#pragma unroll
  for (int i = 0; i < 4; ++i) d[((lane >> 4) + 4 * i) * 16 + (lane & 15)] = acc[i];
  __syncthreads();
  for (int t = threadIdx.x; t < n * n; t += blockDim.x) {
    const int i = t / n, j = t - i * n;
    double s = 0.0;
    for (int w = 0; w < nw; ++w) {
      const double* dw = lds_d + w * 256;
      s += dw[i * 16 + j];
      if (GROUPS == 2) s += dw[(8 + i) * 16 + 8 + j];
    }
    st_ag(part + static_cast<size_t>(t) * nblocks + block, s);
  }
}

template <int GROUPS>
__device__ __forceinline__ void multi_gram_body(const MultiArgs a, int depth, const float* __restrict__ bv,
                 const float* __restrict__ cash, const double* __restrict__ Y,
                 const unsigned char* __restrict__ mem, const DepthState* __restrict__ state,
                 const double* __restrict__ stats_part, double* __restrict__ norm,
                 double* __restrict__ means_out, double* __restrict__ sds_out,
                 double* __restrict__ gram_part, int blk, int nblk) {
  __shared__ double lds_d[(kMultiThreads / 64) * 256];
  __shared__ double sums[2 * kMaxBasisVars];
  __shared__ double msd[2 * kMaxBasisVars];
  if (!ld_ag(&state[depth].active)) return;
  const double n_pre = static_cast<double>(ld_ag(&state[depth].n_pre));
  if (a.normalise && a.v > 0) reduce_partials<true>(stats_part, nblk, 2 * a.v, sums);
  if (threadIdx.x < kMaxBasisVars) {
    const int i = threadIdx.x;
    double m = 0.0, sd = 1.0;
    if (a.normalise && i < a.v) {
      const double s1 = sums[2 * i] / n_pre, s2 = sums[2 * i + 1] / n_pre;
      m = s1;
      sd = sqrt(s2 - s1 * s1);
    }
    msd[i] = m;
    msd[kMaxBasisVars + i] = sd;
    if (blk == 0 && i < a.v) {
      st_ag(norm + norm_at(a.e, a.R, depth, 0, a.max_vars) + i, m);
      st_ag(norm + norm_at(a.e, a.R, depth, 1, a.max_vars) + i, sd);
      const size_t o = (static_cast<size_t>(a.e) * a.R + depth) * a.max_vars + i;
      if (means_out) st_ag(means_out + o, m);
      if (sds_out) st_ag(sds_out + o, sd);
    }
  }
  __syncthreads();
  constexpr int kPathsPerOp = 4 * GROUPS;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const size_t P = static_cast<size_t>(a.P);
  const bool cross = a.cross != 0;
  const int n = a.k + 1;
  const int j = GROUPS == 2 ? (lane & 7) : (lane & 15);
  const int sub = GROUPS == 2 ? ((lane >> 3) & 1) * 4 + (lane >> 4) : (lane >> 4);
  doublex4 acc = {0.0, 0.0, 0.0, 0.0};
  const int stride = nblk * nw * 64;
  for (int base = (blk * nw + wave) * 64; base < a.P; base += stride) {
    const int p = base + lane;
    double x[kMaxBasisVars] = {0.0, 0.0, 0.0, 0.0};
    double t = 0.0, w = 0.0;
    if (p < a.P && mem[p]) {
      norm_vars(bv, a.e, a.max_vars, P, p, a.v, msd, msd + kMaxBasisVars, x);
      t = static_cast<double>(cash[static_cast<size_t>(a.e) * P + p]) + Y[p];
      w = 1.0;
    }
    // 64 paths of this wave -> 64 / (4 GROUPS) MFMAs; lane l takes row j of path `sub`.
    for (int q = 0; q < 64 / kPathsPerOp; ++q) {
      const int srcl = q * kPathsPerOp + sub;
      const double s0 = __shfl(x[0], srcl, 64), s1 = __shfl(x[1], srcl, 64);
      const double s2 = __shfl(x[2], srcl, 64), s3 = __shfl(x[3], srcl, 64);
      const double st = __shfl(t, srcl, 64), sw = __shfl(w, srcl, 64);
      double val = j < a.k ? basis_value(j, a.v, cross, s0, s1, s2, s3) : (j == a.k ? st : 0.0);
      val *= sw;
      acc = mfma16x16x4_f64(val, val, acc);
    }
  }
  block_gram_f64<GROUPS>(acc, n, lds_d, gram_part, blk, nblk);
}

template <int GROUPS>
__global__ void __launch_bounds__(kMultiThreads)
    k_multi_gram(const MultiArgs a, int depth, const float* __restrict__ bv,
                 const float* __restrict__ cash, const double* __restrict__ Y,
                 const unsigned char* __restrict__ mem, const DepthState* __restrict__ state,
                 const double* __restrict__ stats_part, double* __restrict__ norm,
                 double* __restrict__ means_out, double* __restrict__ sds_out,
                 double* __restrict__ gram_part) {
  multi_gram_body<GROUPS>(a, depth, bv, cash, Y, mem, state, stats_part, norm, means_out, sds_out, gram_part, blockIdx.x, gridDim.x);
}


// ---------------------------------------------------------------------------
// Normal-equation solve and net continuation values.
// ---------------------------------------------------------------------------

// The normal equations M beta = t with M = G[0..k)[0..k) / n, t = G[.][k] / n
// (G: (k+1) x (k+1) Gram of basis and target).  Cholesky first: when it
// succeeds and 1 / trace(M^-1) >= tol * trace(M) -- a lower bound of the
// smallest eigenvalue against an upper bound of the largest -- no eigenvalue
// is below tol times the largest, the truncated solve keeps every direction
// and equals the Cholesky solution.  Otherwise the truncated Jacobi solve.
// Single lane; A, V are 16 x 16 LDS scratch, b and z 16.
// The Cholesky branch of solve_multi with the factor in registers (K known
// at compile time): the same operations in the same order (divisions by the
// pivots, not reciprocals), so the same bits.  Returns false where
// solve_multi goes on to the truncated eigen-solve.
template <int K>
__device__ bool cholesky_multi_reg(const double* G, int ld, double inv, double* beta) {
  double L[K][K];
  double trM = 0.0;
#pragma unroll
  for (int i = 0; i < K; ++i) {
    trM += G[i * ld + i] * inv;
#pragma unroll
    for (int j = 0; j <= i; ++j) {
      double s = G[i * ld + j] * inv;
#pragma unroll
      for (int q = 0; q < j; ++q) s -= L[i][q] * L[j][q];
      if (i == j) {
        if (!(s > 0.0)) return false;
        L[i][i] = sqrt(s);
      } else {
        L[i][j] = s / L[j][j];
      }
    }
  }
  // trace(M^-1) = ||L^-1||_F^2, column by column of L^-1
  double trInv = 0.0;
#pragma unroll
  for (int c = 0; c < K; ++c) {
    double z[K];
#pragma unroll
    for (int i = c; i < K; ++i) {
      double s = i == c ? 1.0 : 0.0;
#pragma unroll
      for (int q = c; q < i; ++q) s -= L[i][q] * z[q];
      z[i] = s / L[i][i];
      trInv += z[i] * z[i];
    }
  }
  if (!(trInv * kSvdTolerance * trM <= 1.0)) return false;
  double z[K], x[K];
#pragma unroll
  for (int i = 0; i < K; ++i) {
    double s = G[i * ld + K] * inv;
#pragma unroll
    for (int q = 0; q < i; ++q) s -= L[i][q] * z[q];
    z[i] = s / L[i][i];
  }
#pragma unroll
  for (int i = K - 1; i >= 0; --i) {
    double s = z[i];
#pragma unroll
    for (int q = i + 1; q < K; ++q) s -= L[q][i] * x[q];
    x[i] = s / L[i][i];
  }
#pragma unroll
  for (int i = 0; i < K; ++i) beta[i] = x[i];
  return true;
}

__device__ void solve_multi(const double* G, int k, double n, double* beta, double* A, double* V,
                            double* b, double* z) {
  const int ld = k + 1;
  const double inv = 1.0 / n;
  // the basis sizes of one or two variables and three plain: the Cholesky
  // branch in registers, else straight to the eigen-solve below
  bool reg = true, done = false;
  switch (k) {
    case 1: done = cholesky_multi_reg<1>(G, ld, inv, beta); break;
    case 3: done = cholesky_multi_reg<3>(G, ld, inv, beta); break;
    case 5: done = cholesky_multi_reg<5>(G, ld, inv, beta); break;
    case 6: done = cholesky_multi_reg<6>(G, ld, inv, beta); break;
    case 7: done = cholesky_multi_reg<7>(G, ld, inv, beta); break;
    default: reg = false; break;
  }
  if (done) return;
  for (int i = 0; i < k; ++i) {
    for (int j = 0; j < k; ++j) A[i * 16 + j] = G[i * ld + j] * inv;
    b[i] = G[i * ld + k] * inv;
  }
  if (reg) {
    jacobi_truncated(k, A, V, b, kSvdTolerance, beta);
    return;
  }
  // Cholesky into V (lower triangle)
  bool ok = true;
  double trM = 0.0;
  for (int i = 0; i < k && ok; ++i) {
    trM += A[i * 16 + i];
    for (int j = 0; j <= i; ++j) {
      double s = A[i * 16 + j];
      for (int q = 0; q < j; ++q) s -= V[i * 16 + q] * V[j * 16 + q];
      if (i == j) {
        if (!(s > 0.0)) {
          ok = false;
          break;
        }
        V[i * 16 + i] = sqrt(s);
      } else {
        V[i * 16 + j] = s / V[j * 16 + j];
      }
    }
  }
  if (ok) {
    // trace(M^-1) = ||L^-1||_F^2, column by column of L^-1
    double trInv = 0.0;
    for (int c = 0; c < k; ++c) {
      for (int i = c; i < k; ++i) {
        double s = i == c ? 1.0 : 0.0;
        for (int q = c; q < i; ++q) s -= V[i * 16 + q] * z[q];
        z[i] = s / V[i * 16 + i];
        trInv += z[i] * z[i];
      }
    }
    if (trInv * kSvdTolerance * trM <= 1.0) {
      for (int i = 0; i < k; ++i) {
        double s = b[i];
        for (int q = 0; q < i; ++q) s -= V[i * 16 + q] * z[q];
        z[i] = s / V[i * 16 + i];
      }
      for (int i = k - 1; i >= 0; --i) {
        double s = z[i];
        for (int q = i + 1; q < k; ++q) s -= V[q * 16 + i] * beta[q];
        beta[i] = s / V[i * 16 + i];
      }
      return;
    }
  }
  jacobi_truncated(k, A, V, b, kSvdTolerance, beta);
}

__device__ __forceinline__ void multi_net_body(const MultiArgs a, int depth, const float* __restrict__ bv,
                const float* __restrict__ ev, const unsigned char* __restrict__ mem,
                const DepthState* __restrict__ state, const double* __restrict__ norm,
                const double* __restrict__ gram_part, double* __restrict__ coeffs,
                unsigned long long* __restrict__ keys, double* __restrict__ net_part, int blk, int nblk) {
  __shared__ double gram[kMaxGram * kMaxGram];
  __shared__ double A[16 * 16], V[16 * 16], rhs[16], tmp[16];
  __shared__ double beta[kMaxGram];
  __shared__ double msd[2 * kMaxBasisVars];
  __shared__ double red[16 * 8];
  if (!ld_ag(&state[depth].active)) return;
  const int n = a.k + 1;
  const size_t P = static_cast<size_t>(a.P);
  // the lane's first path's inputs are loaded before the Gram sums and the
  // solve, so their latency overlaps them
  const int p0 = blk * blockDim.x + threadIdx.x;
  unsigned char m0 = 0;
  float raw0[kMaxBasisVars], ev0 = 0.0f;
#pragma unroll
  for (int i = 0; i < kMaxBasisVars; ++i) raw0[i] = 0.0f;
  if (p0 < a.P) {
    m0 = mem[p0];
#pragma unroll
    for (int i = 0; i < kMaxBasisVars; ++i)
      if (i < a.v) raw0[i] = bv[(static_cast<size_t>(a.e) * a.max_vars + i) * P + p0];
    ev0 = ev[static_cast<size_t>(a.e) * P + p0];
  }
  reduce_partials<true>(gram_part, nblk, n * n, gram);
  if (threadIdx.x == 0) {
    solve_multi(gram, a.k, static_cast<double>(ld_ag(&state[depth].n_pre)), beta, A, V, rhs, tmp);
    if (blk == 0)
      for (int i = 0; i < a.maxd; ++i)
        st_ag(coeffs + (static_cast<size_t>(a.e) * a.R + depth) * a.maxd + i,
              i < a.k ? beta[i] : 0.0);
  }
  // the normalisation on the second wavefront, beside the solve on the first
  if (threadIdx.x >= 64 && threadIdx.x < 64 + 2 * kMaxBasisVars) {
    const int t = threadIdx.x - 64;
    const int which = t / kMaxBasisVars, i = t % kMaxBasisVars;
    msd[t] = i < a.v ? ld_ag(norm + norm_at(a.e, a.R, depth, which, a.max_vars) + i)
                     : (which ? 1.0 : 0.0);
  }
  __syncthreads();
  double s[2] = {0.0, 0.0};
  for (int p = p0; p < a.P; p += nblk * blockDim.x) {
    unsigned long long key = kKeySentinel;
    const bool first = p == p0;
    if (first ? m0 != 0 : mem[p] != 0) {
      double x[kMaxBasisVars];
      if (first) {
#pragma unroll
        for (int i = 0; i < kMaxBasisVars; ++i)
          x[i] = i < a.v ? (static_cast<double>(raw0[i]) - msd[i]) / msd[kMaxBasisVars + i] : 0.0;
      } else {
        norm_vars(bv, a.e, a.max_vars, P, p, a.v, msd, msd + kMaxBasisVars, x);
      }
      const double net = basis_dot(a.v, a.k, a.cross != 0, x, beta) -
                         static_cast<double>(first ? ev0 : ev[static_cast<size_t>(a.e) * P + p]);
      s[0] += net;
      s[1] += net * net;
      key = static_cast<unsigned long long>(__double_as_longlong(fabs(net)));
    }
    st_ag(keys + p, key);
  }
  block_sums(s, 2, red, net_part, blk, nblk);
}

__global__ void __launch_bounds__(kMultiThreads)
    k_multi_net(const MultiArgs a, int depth, const float* __restrict__ bv,
                const float* __restrict__ ev, const unsigned char* __restrict__ mem,
                const DepthState* __restrict__ state, const double* __restrict__ norm,
                const double* __restrict__ gram_part, double* __restrict__ coeffs,
                unsigned long long* __restrict__ keys, double* __restrict__ net_part) {
  multi_net_body(a, depth, bv, ev, mem, state, norm, gram_part, coeffs, keys, net_part, blockIdx.x, gridDim.x);
}


// ---------------------------------------------------------------------------
// The selector's band search with exact counts, without ordering the keys.
//
// The gold's search asks for counts of members with |net| <= g at a ladder
// of cuts fixed by the net values' standard deviation -- the growth steps
// G_j = G_0 / m^j, the shrink steps G_0 m^(i+1) and the first shrink step
// G_j m after j growth steps -- and then bisects between two of them.
// k_multi_ladder histograms every member key against the whole ladder (one
// pass), k_multi_select replays the search on those exact counts, and when
// it ends in a bisection, k_multi_tree histograms the keys against every
// guess of the first bisection steps and k_multi_bisect replays those.  A search that runs
// past the ladder (kLadder steps either way) continues in k_multi_select with
// block-wide counts over all keys, so every count the gold would take is
// taken exactly.
// ---------------------------------------------------------------------------
constexpr int kLadder = 64;                   // growth / shrink steps held
constexpr int kLadderSlots = 256;             // >= 3 kLadder - 1
static_assert(3 * kLadder - 1 <= kLadderSlots, "ladder slots");
// ladder_build uses lanes 0 and 64; block sums use one LDS slot per wavefront
static_assert(kSearchThreads >= 256 && kMultiThreads >= 256,
              "ladder needs 128 lanes, the selection's sums 4 wavefronts");
static_assert(kSearchThreads <= 1024 && kMultiThreads <= 1024, "16 wavefront slots");
constexpr unsigned long long kEmpty = ~0ull;  // a cut that counts no path

// Bisection hand-over from k_multi_select to k_multi_tree / k_multi_bisect.
struct SelectState {
  double low, high;
  long long count_low;
  int bisect;  // 1: the search continues by bisection in (low, high]
  int it;      // the gold's bisection step counter so far
};

// Extent of the member keys: smallest positive key, largest non-NaN key
// (+inf included), count of zero keys and of non-NaN keys.  Every count at a
// cut below the smallest positive key is the zero count, at or above the
// largest key the non-NaN count, so the search's long tails (e.g. half the
// net values exactly 0, where the gold's shrink loop runs all its 1000
// steps) cost no pass over the keys.
struct KeyExtent {
  unsigned long long minpos, maxval;
  long long nzero, nvalid;
};

constexpr unsigned long long kInfKey = 0x7FF0000000000000ull;

// Wavefront `wave` of the block sums the per-block extents ext[0..nb) into
// LDS *out (its lanes load four blocks' extents at a time, all in flight);
// the other wavefronts return at once.  No barrier.
__device__ void reduce_extents_wave(const KeyExtent* __restrict__ ext, int nb, KeyExtent* out,
                                    int wave) {
  if ((threadIdx.x >> 6) != wave) return;
  const int lane = threadIdx.x & 63;
  unsigned long long mn = ~0ull, mx = 0ull;
  long long nz = 0, nv = 0;
  for (int b0 = 0; b0 < nb; b0 += 256) {
    unsigned long long emn[4], emx[4];
    long long enz[4], env[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int b = b0 + lane + 64 * r;
      const bool ok = b < nb;
      emn[r] = ok ? ld_ag(&ext[b].minpos) : ~0ull;
      emx[r] = ok ? ld_ag(&ext[b].maxval) : 0ull;
      enz[r] = ok ? ld_ag(&ext[b].nzero) : 0ll;
      env[r] = ok ? ld_ag(&ext[b].nvalid) : 0ll;
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      mn = emn[r] < mn ? emn[r] : mn;
      mx = emx[r] > mx ? emx[r] : mx;
      nz += enz[r];
      nv += env[r];
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long a = __shfl_xor(mn, o, 64), b = __shfl_xor(mx, o, 64);
    mn = a < mn ? a : mn;
    mx = b > mx ? b : mx;
    nz += __shfl_xor(nz, o, 64);
    nv += __shfl_xor(nv, o, 64);
  }
  if (lane == 0) *out = KeyExtent{mn, mx, nz, nv};
}

// The same on wavefront 0, then a barrier.  Whole block.
__device__ void reduce_extents(const KeyExtent* __restrict__ ext, int nb, KeyExtent* out) {
  reduce_extents_wave(ext, nb, out, 0);
  __syncthreads();
}

__device__ __forceinline__ unsigned long long key_of(double g) {
  return static_cast<unsigned long long>(__double_as_longlong(g + 0.0));
}

// Net-value mean and standard deviation as the gold forms them.
__device__ __forceinline__ double net_sd(const double* sums, long long n_pre) {
  const double np = static_cast<double>(n_pre);
  const double mean = sums[0] / np;
  return sqrt(sums[1] / np - mean * mean);
}

// Number of ladder values ladder_build makes for n_pre members: it depends
// on the selector and the member count only, not on the net values.
__device__ __forceinline__ int ladder_count(const MultiArgs& a, long long n_pre) {
  const double np = static_cast<double>(n_pre);
  const long long lowerT = static_cast<long long>(np * a.lower_frac);
  const long long upperT = static_cast<long long>(np * a.upper_frac);
  if (a.selector == 1 || (n_pre > lowerT && n_pre < upperT)) return 1;
  return 3 * kLadder - 1;
}

// The ladder of raw cut values v[0..n) and their key bounds tb (a member
// counts at cut t when key <= tb[t]; kEmpty counts none).  Layout: fraction
// selector with the pre-selected count inside the window: v[0] = 10000;
// fraction selector otherwise: v[j] = G_j (j < L), v[L + i] = G_0 m^(i+1)
// (i < L, by repeated multiplication as the gold's shrink loop), v[2L + j -
// 1] = G_j m (1 <= j < L); standard-deviation selector: v[0] = sd_guess sd
// with the open bound.  Whole block; returns n.
// The growth steps v[0..L) (G_0 divided by m again and again) and the shrink
// steps v[L..2L) (G_0 m multiplied by m again and again), as ladder_build
// forms them.  One lane; sums in LDS.
__device__ void ladder_chains(const MultiArgs& a, long long n_pre, const double* sums, double* v) {
  const int L = kLadder;
  const double sd = net_sd(sums, n_pre);
  double g = sd * a.sd_guess;
  v[0] = g;
  for (int j = 1; j < L; ++j) {
    g /= a.multiplier;
    v[j] = g;
  }
  double lo = (sd * a.sd_guess) * a.multiplier;
  v[L] = lo;
  for (int i = 1; i < L; ++i) {
    lo *= a.multiplier;
    v[L + i] = lo;
  }
}

// chains_done: ladder_chains has already filled v[0..2L) (the full ladder
// case only).
__device__ int ladder_build(const MultiArgs& a, long long n_pre, const double* sums, double* v,
                            unsigned long long* tb, bool chains_done = false) {
  const int t = threadIdx.x, L = kLadder;
  const double sd = net_sd(sums, n_pre);
  const double np = static_cast<double>(n_pre);
  const long long lowerT = static_cast<long long>(np * a.lower_frac);
  const long long upperT = static_cast<long long>(np * a.upper_frac);
  int n;
  if (a.selector == 1) {
    n = 1;
    if (t == 0) {
      v[0] = a.sd_guess * sd;
      tb[0] = v[0] > 0.0 ? static_cast<unsigned long long>(__double_as_longlong(v[0])) - 1ull
                         : kEmpty;
    }
  } else if (n_pre > lowerT && n_pre < upperT) {
    n = 1;
    if (t == 0) {
      v[0] = 10000.0;
      tb[0] = key_of(10000.0);
    }
  } else {
    n = 3 * L - 1;
    if (chains_done) {
    } else if (t == 0) {
      double g = sd * a.sd_guess;
      v[0] = g;
      for (int j = 1; j < L; ++j) {
        g /= a.multiplier;
        v[j] = g;
      }
    } else if (t == 64) {  // the kernels run at least 128 lanes
      double lo = (sd * a.sd_guess) * a.multiplier;
      v[L] = lo;
      for (int i = 1; i < L; ++i) {
        lo *= a.multiplier;
        v[L + i] = lo;
      }
    }
    __syncthreads();
    for (int j = 1 + t; j < L; j += blockDim.x) v[2 * L + j - 1] = v[j] * a.multiplier;
    __syncthreads();
    for (int u = t; u < n; u += blockDim.x) tb[u] = v[u] >= 0.0 ? key_of(v[u]) : kEmpty;
  }
  __syncthreads();
  return n;
}

// Stable ranks of the key bounds: ts (ascending) and rank[t] = position of
// tb[t] in ts.  Whole block.
__device__ __forceinline__ void ladder_sort(int n, const unsigned long long* tb,
                                            unsigned long long* ts, int* rank) {
  for (int t = threadIdx.x; t < n; t += blockDim.x) {
    const unsigned long long x = tb[t];
    int r = 0;
#pragma unroll 8
    for (int u = 0; u < n; ++u) r += (tb[u] < x || (tb[u] == x && u < t)) ? 1 : 0;
    ts[r] = x;
    if (rank) rank[t] = r;
  }
  __syncthreads();
}

// First k < n with ts[k] >= key (n when none).
__device__ __forceinline__ int ladder_bin(const unsigned long long* ts, int n,
                                          unsigned long long key) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (ts[mid] < key)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

// Histogram of the member keys over the ladder: hist_part[bin * nb + block]
// (FP64, exact) counts the members whose first bound >= key is bin.
__device__ __forceinline__ void multi_ladder_body(const MultiArgs a, int depth, const unsigned long long* __restrict__ keys,
                   const double* __restrict__ net_part, int nb_net,
                   const DepthState* __restrict__ state, double* __restrict__ hist_part,
                   KeyExtent* __restrict__ ext_part, int blk, int nblk) {
  __shared__ double sums[2];
  __shared__ double v[kLadderSlots];
  __shared__ unsigned long long tb[kLadderSlots], ts[kLadderSlots];
  __shared__ int hist[kLadderSlots];
  __shared__ KeyExtent wext[16];
  if (!ld_ag(&state[depth].active)) return;
  // the lane's first key is loaded before the ladder is built, so its
  // latency overlaps the net-value sums
  const int p0 = blk * blockDim.x + threadIdx.x;
  const unsigned long long key0 = p0 < a.P ? ld_ag(keys + p0) : kKeySentinel;
  reduce_partials<true>(net_part, nb_net, 2, sums);
  const int n = ladder_build(a, ld_ag(&state[depth].n_pre), sums, v, tb);
  ladder_sort(n, tb, ts, nullptr);
  for (int t = threadIdx.x; t < n; t += blockDim.x) hist[t] = 0;
  __syncthreads();
  unsigned long long mn = ~0ull, mx = 0ull;
  long long nz = 0, nv = 0;
  for (int p = p0; p < a.P; p += nblk * blockDim.x) {
    const unsigned long long key = p == p0 ? key0 : ld_ag(keys + p);
    if (key != kKeySentinel) {
      const int b = ladder_bin(ts, n, key);
      if (b < n) atomicAdd(&hist[b], 1);
      if (key <= kInfKey) {  // not NaN
        ++nv;
        mx = key > mx ? key : mx;
        if (key == 0ull)
          ++nz;
        else
          mn = key < mn ? key : mn;
      }
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long a2 = __shfl_xor(mn, o, 64), b2 = __shfl_xor(mx, o, 64);
    mn = a2 < mn ? a2 : mn;
    mx = b2 > mx ? b2 : mx;
    nz += __shfl_xor(nz, o, 64);
    nv += __shfl_xor(nv, o, 64);
  }
  if ((threadIdx.x & 63) == 0) wext[threadIdx.x >> 6] = KeyExtent{mn, mx, nz, nv};
  __syncthreads();
  for (int t = threadIdx.x; t < n; t += blockDim.x)
    st_ag(hist_part + static_cast<size_t>(t) * nblk + blk, static_cast<double>(hist[t]));
  if (threadIdx.x == 0) {
    KeyExtent e = wext[0];
    for (int w = 1; w < static_cast<int>(blockDim.x >> 6); ++w) {
      e.minpos = wext[w].minpos < e.minpos ? wext[w].minpos : e.minpos;
      e.maxval = wext[w].maxval > e.maxval ? wext[w].maxval : e.maxval;
      e.nzero += wext[w].nzero;
      e.nvalid += wext[w].nvalid;
    }
    st_ag(&ext_part[blk].minpos, e.minpos);
    st_ag(&ext_part[blk].maxval, e.maxval);
    st_ag(&ext_part[blk].nzero, e.nzero);
    st_ag(&ext_part[blk].nvalid, e.nvalid);
  }
}

__global__ void __launch_bounds__(kMultiThreads)
    k_multi_ladder(const MultiArgs a, int depth, const unsigned long long* __restrict__ keys,
                   const double* __restrict__ net_part, int nb_net,
                   const DepthState* __restrict__ state, double* __restrict__ hist_part,
                   KeyExtent* __restrict__ ext_part) {
  multi_ladder_body(a, depth, keys, net_part, nb_net, state, hist_part, ext_part, blockIdx.x, gridDim.x);
}


// Count of member keys <= key_of(g) for g >= 0 from the extents when they
// decide it, else -1.
__device__ __forceinline__ long long extent_count(const KeyExtent& x, double g) {
  const unsigned long long gb = key_of(g);
  if (gb < x.minpos) return x.nzero;
  if (gb >= x.maxval) return x.nvalid;
  return -1;
}

// Block-wide count of the keys[0..n) that are <= gb and not the sentinel.
// Whole block (barriers); uniform result.
__device__ long long block_count(const unsigned long long* __restrict__ keys, long long n,
                                 unsigned long long gb, long long* red /* [16] */) {
  long long c = 0;
  for (long long i = threadIdx.x; i < n; i += blockDim.x) {
    const unsigned long long k = ld_ag(keys + i);
    c += (k != kKeySentinel && k <= gb) ? 1 : 0;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = c;
  __syncthreads();
  long long tot = 0;
  for (int w = 0; w < static_cast<int>(blockDim.x >> 6); ++w) tot += red[w];
  __syncthreads();
  return tot;
}

// Writes the regression's cuts (when more than the cutoff remain) and the
// next depth's members.  Single thread.
__device__ __forceinline__ void finish_select(const MultiArgs& a, int depth, double cut,
                                              long long after, DepthState* state, double* lower,
                                              double* upper) {
  if (after > a.cutoff) {
    st_ag(lower + static_cast<size_t>(a.e) * a.R + depth, -cut);
    st_ag(upper + static_cast<size_t>(a.e) * a.R + depth, cut);
  }
  if (depth + 1 < a.R) {
    st_ag(&state[depth + 1].active, after > a.cutoff ? 1 : 0);
    st_ag(&state[depth + 1].n_pre, after);
    if (a.selector == 1)
      st_ag(&state[depth + 1].cutbits,
            cut > 0.0 ? static_cast<unsigned long long>(__double_as_longlong(cut)) - 1ull : 0ull);
    else
      st_ag(&state[depth + 1].cutbits, cut >= 0.0 ? key_of(cut) : 0ull);
  }
}

// The band search on the ladder counts (RegressionSelector_gold_Fraction::Select,
// Regression_Selector_concrete_gold.cpp:13-183; RegressionSelectorStandardDeviations,
// RegressionSelectorConcrete.cu:19-90).  One block; every thread runs the same
// search on LDS values so the block-wide fallback counts stay uniform.
__device__ __forceinline__ void multi_select_body(const MultiArgs a, int depth, const unsigned long long* __restrict__ keys,
                   const double* __restrict__ net_part, int nb,
                   const double* __restrict__ hist_part, const KeyExtent* __restrict__ ext_part,
                   int nh, DepthState* __restrict__ state, SelectState* __restrict__ sel,
                   double* __restrict__ lower, double* __restrict__ upper, int blk, int nblk) {
  __shared__ double sums[2];
  __shared__ double v[kLadderSlots], cnt[kLadderSlots];
  __shared__ unsigned long long tb[kLadderSlots], ts[kLadderSlots];
  __shared__ int rank[kLadderSlots];
  __shared__ long long red[16];
  __shared__ KeyExtent ext;
  if (!ld_ag(&state[depth].active)) {
    if (threadIdx.x == 0) {
      st_ag(&sel->bisect, 0);
      if (depth + 1 < a.R) st_ag(&state[depth + 1].active, 0);
    }
    return;
  }
  const long long n_pre = ld_ag(&state[depth].n_pre);
  // the three cross-block sums at once, on disjoint wavefronts: the ladder's
  // bin counts (cnt[k]: members in bin k), the net-value sums, the extents
  // (the net-value wavefront then builds the ladder's two serial chains
  // while the bin counts are still being summed)
  const int nw = blockDim.x >> 6;
  const int nlad = ladder_count(a, n_pre);
  reduce_partials_waves<true>(hist_part, nh, nlad, cnt, 0, nw - 2);
  reduce_partials_waves<true>(net_part, nb, 2, sums, nw - 2, 1);
  if (nlad > 1 && threadIdx.x == (nw - 2) * 64) ladder_chains(a, n_pre, sums, v);
  reduce_extents_wave(ext_part, nh, &ext, nw - 1);
  __syncthreads();
  const int n = ladder_build(a, n_pre, sums, v, tb, nlad > 1);
  ladder_sort(n, tb, ts, rank);
  if (threadIdx.x == 0) {
    // running count in a register, so the bins' loads are independent of
    // the additions (the counts are integers, exact in FP64)
    double run = cnt[0];
#pragma unroll 8
    for (int k = 1; k < n; ++k) {
      run += cnt[k];
      cnt[k] = run;
    }
  }
  __syncthreads();
  // exact count at ladder value t
  auto C = [&](int t) -> long long {
    return tb[t] == kEmpty ? 0 : static_cast<long long>(cnt[rank[t]]);
  };
  auto brute = [&](double g) -> long long {
    if (!(g >= 0.0)) return 0;
    const long long c = extent_count(ext, g);
    return c >= 0 ? c : block_count(keys, a.P, key_of(g), red);
  };
  const double np = static_cast<double>(n_pre);
  const long long lowerT = static_cast<long long>(np * a.lower_frac);
  const long long upperT = static_cast<long long>(np * a.upper_frac);
  const double m = a.multiplier;
  const int L = kLadder;
  double cut;
  long long after;
  int bisect = 0;
  double low = 0.0, high = 0.0;
  if (a.selector == 1 || (n_pre > lowerT && n_pre < upperT)) {
    cut = v[0];
    after = C(0);
  } else {
    double guess = v[0];
    long long count = C(0);
    int it = 0, j = 0;
    const int lane = threadIdx.x & 63;
    static_assert(kLadder == 64, "one ladder step per lane");
    if (count < lowerT) {
      // the growth loop's steps inside the ladder at once: it stops at the
      // first j >= 1 whose count reaches lowerT (one ballot per wavefront,
      // the same in every wavefront), else it has walked the whole ladder
      const unsigned long long hit = __ballot(lane >= 1 && C(lane) >= lowerT);
      j = hit ? __builtin_ctzll(hit) : L - 1;
      it = j;
      guess = v[j];
      count = C(j);
    }
    while (count < lowerT && it++ < kSelectMaxSteps) {
      ++j;
      if (j < L) {
        guess = v[j];
        count = C(j);
      } else {
        guess /= m;
        if (guess >= 0.0 && key_of(guess) >= ext.maxval && ext.nvalid < lowerT) {
          // every later guess is larger, so the count stays at nvalid: the
          // gold's loop runs out its steps (the same divisions, no counts)
          while (it++ < kSelectMaxSteps) guess /= m;
          count = ext.nvalid;
          break;
        }
        count = brute(guess);
      }
    }
    if (count < upperT) {
      cut = guess;
      after = count;
    } else {
      if (j == 0) {
        low = v[L];
        count = C(L);
      } else if (j < L) {
        low = v[2 * L + j - 1];
        count = C(2 * L + j - 1);
      } else {
        low = guess * m;
        count = brute(low);
      }
      it = 0;
      int i = 0;
      if (j == 0 && count > upperT) {
        // the shrink loop's steps inside the ladder at once, as above: the
        // first i >= 1 whose count falls to upperT, else the whole ladder
        const unsigned long long hit = __ballot(lane >= 1 && C(L + lane) <= upperT);
        i = hit ? __builtin_ctzll(hit) : L - 1;
        it = i;
        low = v[L + i];
        count = C(L + i);
      }
      while (count > upperT && it++ < kSelectMaxSteps) {
        ++i;
        if (j == 0 && i < L) {
          low = v[L + i];
          count = C(L + i);
        } else {
          low *= m;
          if (low >= 0.0 && key_of(low) < ext.minpos && ext.nzero > upperT) {
            // every later cut is smaller, so the count stays at the zero
            // count: the gold's loop runs out its steps (same products)
            while (it++ < kSelectMaxSteps) low *= m;
            count = ext.nzero;
            break;
          }
          count = brute(low);
        }
      }
      if (count > lowerT) {
        cut = low;
        after = count;
      } else {
        bisect = 1;
        high = guess;
        cut = 0.0;
        after = count;  // count at low, handed to the bisection
      }
    }
  }
  if (threadIdx.x == 0) {
    st_ag(&sel->bisect, bisect);
    if (bisect) {
      st_ag(&sel->low, low);
      st_ag(&sel->high, high);
      st_ag(&sel->count_low, after);
      st_ag(&sel->it, 0);
    } else {
      finish_select(a, depth, cut, after, state, lower, upper);
    }
  }
}

__global__ void __launch_bounds__(kSearchThreads)
    k_multi_select(const MultiArgs a, int depth, const unsigned long long* __restrict__ keys,
                   const double* __restrict__ net_part, int nb,
                   const double* __restrict__ hist_part, const KeyExtent* __restrict__ ext_part,
                   int nh, DepthState* __restrict__ state, SelectState* __restrict__ sel,
                   double* __restrict__ lower, double* __restrict__ upper) {
  multi_select_body(a, depth, keys, net_part, nb, hist_part, ext_part, nh, state, sel, lower, upper, blockIdx.x, gridDim.x);
}


// The bisection tree: node 1 is the first guess 0.5 (low + high); node i
// (heap order) splits its interval at its guess, its left child 2i takes the
// lower half (the gold sets high = guess), its right child 2i + 1 the upper
// half (low = guess).  The guesses of a depth-kTreeDepth tree cover every
// guess the first kTreeDepth bisection steps can take, with the same
// arithmetic as the gold's loop; in-order position of node i at level k is
// (2 (i - 2^k) + 1) 2^(D-1-k) - 1, and the guesses ascend in that order.
constexpr int kTreeDepth = 10;
constexpr int kTreeNodes = (1 << kTreeDepth) - 1;
constexpr int kBisectStages = 2;  // bisection steps held: kBisectStages * kTreeDepth
constexpr int kHistBlocks = 256;  // blocks of k_multi_ladder and k_multi_tree (at most)

__device__ __forceinline__ int tree_rank(int i) {
  const int k = 31 - __builtin_clz(i);  // level
  return (2 * (i - (1 << k)) + 1) * (1 << (kTreeDepth - 1 - k)) - 1;
}

// Per block: cumulative counts, in in-order position, of the member keys in
// (low, high] that are <= each tree guess: cum_part[rank * nb + block].
__device__ __forceinline__ void multi_tree_body(const MultiArgs a, int depth, const unsigned long long* __restrict__ keys,
                 const DepthState* __restrict__ state, const SelectState* __restrict__ sel,
                 int* __restrict__ cum_part, int blk, int nblk) {
  __shared__ double lo[kTreeNodes + 1], hi[kTreeNodes + 1];
  __shared__ unsigned long long ts[kTreeNodes];
  __shared__ int hist[kTreeNodes + 1];
  if (!ld_ag(&state[depth].active) || !ld_ag(&sel->bisect)) return;
  const double low = ld_ag(&sel->low), high = ld_ag(&sel->high);
  for (int k = 0; k < kTreeDepth; ++k) {
    for (int i = (1 << k) + threadIdx.x; i < (2 << k); i += blockDim.x) {
      double l = low, h = high;
      if (i > 1) {
        const int q = i >> 1;
        const double g = 0.5 * (lo[q] + hi[q]);
        if (i & 1) {
          l = g;
          h = hi[q];
        } else {
          l = lo[q];
          h = g;
        }
      }
      lo[i] = l;
      hi[i] = h;
      ts[tree_rank(i)] = key_of(0.5 * (l + h));
    }
    __syncthreads();
  }
  for (int t = threadIdx.x; t < kTreeNodes; t += blockDim.x) hist[t] = 0;
  __syncthreads();
  const unsigned long long lb = key_of(low), hb = key_of(high);
  for (int p = blk * blockDim.x + threadIdx.x; p < a.P; p += nblk * blockDim.x) {
    const unsigned long long key = ld_ag(keys + p);
    if (key != kKeySentinel && key > lb && key <= hb) {
      const int b = ladder_bin(ts, kTreeNodes, key);
      if (b < kTreeNodes) atomicAdd(&hist[b], 1);
    }
  }
  __syncthreads();
  if (threadIdx.x < 64) {
    // inclusive prefix over the in-order bins on one wavefront: each lane
    // sums a run of kScanRun bins, the lanes' sums are scanned across the
    // wavefront, then each lane writes its run's prefixes (integer counts,
    // so the order of the additions does not matter)
    constexpr int kScanRun = (kTreeNodes + 63) / 64;
    const int lane = threadIdx.x;
    int v[kScanRun];
    int sum = 0;
#pragma unroll
    for (int q = 0; q < kScanRun; ++q) {
      const int t = lane * kScanRun + q;
      v[q] = t < kTreeNodes ? hist[t] : 0;
      sum += v[q];
    }
    int incl = sum;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int y = __shfl_up(incl, o, 64);
      if (lane >= o) incl += y;
    }
    int run = incl - sum;
#pragma unroll
    for (int q = 0; q < kScanRun; ++q) {
      const int t = lane * kScanRun + q;
      run += v[q];
      if (t < kTreeNodes) hist[t] = run;
    }
  }
  __syncthreads();
  for (int t = threadIdx.x; t < kTreeNodes; t += blockDim.x)
    st_ag(cum_part + static_cast<size_t>(t) * nblk + blk, hist[t]);
}

__global__ void __launch_bounds__(kMultiThreads)
    k_multi_tree(const MultiArgs a, int depth, const unsigned long long* __restrict__ keys,
                 const DepthState* __restrict__ state, const SelectState* __restrict__ sel,
                 int* __restrict__ cum_part) {
  multi_tree_body(a, depth, keys, state, sel, cum_part, blockIdx.x, gridDim.x);
}


// Sums over the blocks of the counts of the tree nodes in the kSubtreeLevels
// levels of the subtree rooted at `root` (heap order; nodes past the tree
// are skipped) into sub[local], local = 2^l - 1 + j for node root 2^l + j:
// one wavefront per node, all in flight at once.  Whole block, barrier.
constexpr int kSubtreeLevels = 4;
constexpr int kSubtreeNodes = (1 << kSubtreeLevels) - 1;
static_assert(kSubtreeNodes <= kSearchThreads / 64 && kSubtreeNodes <= kMultiThreads / 64,
              "one wavefront per subtree node");

__device__ void subtree_counts(const int* __restrict__ cum_part, int nb, int root,
                               long long* sub) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();  // every lane has read the previous subtree's counts
  if (w < kSubtreeNodes) {
    const int l = 31 - __builtin_clz(w + 1);   // level below the root
    const int node = (root << l) + (w + 1 - (1 << l));
    long long c = 0;
    if (node <= kTreeNodes) {
      const size_t row = static_cast<size_t>(tree_rank(node)) * nb;
      for (int b = lane; b < nb; b += 64) c += ld_ag(cum_part + row + b);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
    if (lane == 0) sub[w] = c;
  }
  __syncthreads();
}

// The gold's bisection loop on the tree counts.  Count at a guess g in
// [low, high] = count at low + keys in (low, g].  A stage that runs past its
// tree hands (low, high, count at low, step counter) to the next stage
// (k_multi_tree + k_multi_bisect again); the last stage continues with
// block-wide counts over all keys.
__device__ __forceinline__ void multi_bisect_body(const MultiArgs a, int depth, DepthState* __restrict__ state,
                   SelectState* __restrict__ sel, const unsigned long long* __restrict__ keys,
                   const int* __restrict__ cum_part, const KeyExtent* __restrict__ ext_part,
                   int nb, double* __restrict__ lower, double* __restrict__ upper,
                   int last_stage, int blk, int nblk) {
  __shared__ long long red[16];
  __shared__ long long sub[kSubtreeNodes];
  __shared__ KeyExtent ext;
  if (!ld_ag(&state[depth].active) || !ld_ag(&sel->bisect)) return;
  reduce_extents(ext_part, nb, &ext);
  const long long n_pre = ld_ag(&state[depth].n_pre);
  const double np = static_cast<double>(n_pre);
  const long long lowerT = static_cast<long long>(np * a.lower_frac);
  const long long upperT = static_cast<long long>(np * a.upper_frac);
  long long c_low = ld_ag(&sel->count_low);
  double low = ld_ag(&sel->low), high = ld_ag(&sel->high);
  int it = ld_ag(&sel->it);
  int node = 1;
  const long long c0 = c_low;  // the tree counts keys in (low, high] of this stage
  // the counts of a subtree of nodes are summed together, so kSubtreeLevels
  // steps of the walk take one round of cross-block loads
  int sub_root = 0;
  auto count_at = [&](double g) -> long long {
    if (node <= kTreeNodes) {
      const int dl = sub_root ? (31 - __builtin_clz(node)) - (31 - __builtin_clz(sub_root)) : -1;
      if (dl < 0 || dl >= kSubtreeLevels || (node >> dl) != sub_root) {
        sub_root = node;
        subtree_counts(cum_part, nb, sub_root, sub);
        return c0 + sub[0];
      }
      return c0 + sub[(1 << dl) - 1 + (node - (sub_root << dl))];
    }
    if (!(g >= 0.0)) return 0;
    const long long c = extent_count(ext, g);
    return c >= 0 ? c : block_count(keys, a.P, key_of(g), red);
  };
  double guess = 0.5 * (low + high);
  long long count = count_at(guess);
  while ((count < lowerT || count > upperT) && (high - low) > 1e-6 && it++ < kSelectMaxSteps) {
    const bool up = count < lowerT;
    if (up) {
      low = guess;
      c_low = count;
    } else {
      high = guess;
    }
    if (node <= kTreeNodes) node = 2 * node + (up ? 1 : 0);  // past the tree: stays past it
    if (node > kTreeNodes && !last_stage) {
      __syncthreads();
      if (threadIdx.x == 0) {  // hand over; the next stage's root is 0.5 (low + high)
        st_ag(&sel->low, low);
        st_ag(&sel->high, high);
        st_ag(&sel->count_low, c_low);
        st_ag(&sel->it, it);
      }
      return;
    }
    guess = 0.5 * (low + high);
    count = count_at(guess);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    st_ag(&sel->bisect, 0);
    finish_select(a, depth, guess, count, state, lower, upper);
  }
}

__global__ void __launch_bounds__(kSearchThreads)
    k_multi_bisect(const MultiArgs a, int depth, DepthState* __restrict__ state,
                   SelectState* __restrict__ sel, const unsigned long long* __restrict__ keys,
                   const int* __restrict__ cum_part, const KeyExtent* __restrict__ ext_part,
                   int nb, double* __restrict__ lower, double* __restrict__ upper,
                   int last_stage) {
  multi_bisect_body(a, depth, state, sel, keys, cum_part, ext_part, nb, lower, upper, last_stage, blockIdx.x, gridDim.x);
}


// Date 0's decision and sum_p Y_p / B_0 into block partials.
__device__ __forceinline__ void multi_final_body(const MultiArgs a, const float* __restrict__ bv, const float* __restrict__ cash,
                  const float* __restrict__ ev, const float* __restrict__ num,
                  const double* __restrict__ coeffs, const double* __restrict__ lower,
                  const double* __restrict__ upper, const double* __restrict__ norm,
                  const double* __restrict__ Y, double* __restrict__ part, int blk, int nblk) {
  __shared__ double beta[kMaxDepth * kMaxGram];
  __shared__ double nrm[kMaxDepth * 2 * kMaxBasisVars];
  __shared__ double lo[kMaxDepth], hi[kMaxDepth];
  __shared__ double red[16];
  load_decision(a, 0, coeffs, lower, upper, norm, beta, nrm, lo, hi);
  const size_t P = static_cast<size_t>(a.P);
  double s = 0.0;
  for (int p = blk * blockDim.x + threadIdx.x; p < a.P; p += nblk * blockDim.x) {
    const double y = decide(a, 0, a.v, a.k, bv, cash, ev, P, p, Y[p], beta, nrm, lo, hi);
    s += y / static_cast<double>(num[p]);
  }
  block_sum_to<true>(s, red, part + blk);
}

__global__ void __launch_bounds__(kMultiThreads)
    k_multi_final(const MultiArgs a, const float* __restrict__ bv, const float* __restrict__ cash,
                  const float* __restrict__ ev, const float* __restrict__ num,
                  const double* __restrict__ coeffs, const double* __restrict__ lower,
                  const double* __restrict__ upper, const double* __restrict__ norm,
                  const double* __restrict__ Y, double* __restrict__ part) {
  multi_final_body(a, bv, cash, ev, num, coeffs, lower, upper, norm, Y, part, blockIdx.x, gridDim.x);
}


__global__ void k_fill_zero(double* __restrict__ a, size_t n) {
  for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<size_t>(gridDim.x) * blockDim.x)
    a[i] = 0.0;
}

// ---------------------------------------------------------------------------
// The whole pass in one persistent launch: the same phase bodies, one after
// the other, every block of the grid taking part in the multi-block phases
// and block 0 in the one-block searches, with a grid-wide barrier between
// phases (what a phase hands to other blocks is stored and loaded
// agent-coherent, so the barrier waits only for each wave's stores).  Each
// lane's per-path values (running values, membership) stay with the lane
// that wrote them.  Depths a date does not reach
// and bisection stages the search does not need are skipped, uniformly over
// the grid (every block reads the same state after a barrier).  Launched with
// at most as many blocks as can be resident; every wait is bounded, and a
// wait that gives up stops the pass with the error word set and a NaN price.
// ---------------------------------------------------------------------------
struct MultiDates {
  signed char v[kMaxRates];  // variables at each exercise date
};

// Grid-wide barrier without a shared counter (a single word that every
// block adds to serialises hundreds of atomics): block b stores the phase
// number into flags[b] once its waves' stores are performed, and every block
// waits until each of the grid's flags has reached the phase (thread t reads
// flags[t]).  The values blocks exchange go through agent-coherent stores
// and loads, so no cache write-back or invalidation is needed.  Bounded;
// false on every lane if a wait gave up.
__device__ __forceinline__ bool multi_grid_sync(unsigned* flags, unsigned phase) {
  __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): this wave's stores performed
  __syncthreads();
  if (threadIdx.x == 0) st_ag(flags + blockIdx.x, phase);
  int ok = 1;
  if (threadIdx.x < gridDim.x) {
    unsigned polls = 0;
    while (ld_ag(flags + threadIdx.x) < phase) {
      __builtin_amdgcn_s_sleep(1);
      if (++polls > kBarrierPolls) {
        ok = 0;
        break;
      }
    }
  }
  return __syncthreads_and(ok) != 0;
}

__global__ void __launch_bounds__(kMultiThreads)
    k_multi_pass(MultiArgs a, const MultiDates dates, int nb, int nh, const float* __restrict__ bv,
                 const float* __restrict__ cash, const float* __restrict__ ev,
                 const float* __restrict__ num, double* coeffs, double* lower, double* upper,
                 double* norm, double* means, double* sds, double* Y, unsigned char* mem,
                 unsigned long long* keys, int* cum, double* stats, double* gram, double* net,
                 double* hist, KeyExtent* ext, double* price_part, DepthState* state,
                 SelectState* sel, unsigned* bar, double* price_out) {
  const int blk = blockIdx.x, G = gridDim.x;
  const bool cross = a.cross != 0;
  unsigned phase = 0;
  // outputs start at zero (depths never reached stay zero)
  {
    const size_t ncoef = static_cast<size_t>(a.E) * a.R * a.maxd, ncut = static_cast<size_t>(a.E) * a.R;
    const size_t nvar = static_cast<size_t>(a.E) * a.R * a.max_vars;
    const size_t stride = static_cast<size_t>(G) * blockDim.x;
    size_t nmax = 2 * nvar + 1;
    nmax = ncoef > nmax ? ncoef : nmax;
    nmax = ncut > nmax ? ncut : nmax;
    for (size_t i = static_cast<size_t>(blk) * blockDim.x + threadIdx.x; i < nmax; i += stride) {
      if (i < 2 * nvar + 1) st_ag(norm + i, 0.0);
      if (i < ncoef) st_ag(coeffs + i, 0.0);
      if (i < ncut) {
        st_ag(lower + i, 0.0);
        st_ag(upper + i, 0.0);
      }
      if (i < nvar && means) st_ag(means + i, 0.0);
      if (i < nvar && sds) st_ag(sds + i, 0.0);
    }
  }
  if (!multi_grid_sync(bar + 2, ++phase)) goto fail;
  for (int e = a.E - 1; e >= 0; --e) {
    a.e = e;
    a.v = dates.v[e];
    a.k = basis_count(a.v, cross);
    a.vn = e + 1 < a.E ? dates.v[e + 1] : 0;
    a.kn = e + 1 < a.E ? basis_count(a.vn, cross) : 0;
    for (int d = 0; d < a.R; ++d) {
      if (d > 0 && !ld_ag(&state[d].active)) break;  // no deeper regression at this date
      if (blk < nb)
        multi_prep_body(a, d, bv, cash, ev, num, coeffs, lower, upper, norm, Y, mem, keys, state,
                        stats, blk, nb);
      if (!multi_grid_sync(bar + 2, ++phase)) goto fail;
      if (blk < nb) {
        if (a.k + 1 <= 8)
          multi_gram_body<2>(a, d, bv, cash, Y, mem, state, stats, norm, means, sds, gram, blk,
                             nb);
        else
          multi_gram_body<1>(a, d, bv, cash, Y, mem, state, stats, norm, means, sds, gram, blk,
                             nb);
      }
      if (!multi_grid_sync(bar + 2, ++phase)) goto fail;
      if (blk < nb)
        multi_net_body(a, d, bv, ev, mem, state, norm, gram, coeffs, keys, net, blk, nb);
      if (!multi_grid_sync(bar + 2, ++phase)) goto fail;
      if (blk < nh) multi_ladder_body(a, d, keys, net, nb, state, hist, ext, blk, nh);
      if (!multi_grid_sync(bar + 2, ++phase)) goto fail;
      if (blk == 0)
        multi_select_body(a, d, keys, net, nb, hist, ext, nh, state, sel, lower, upper, 0, 1);
      if (!multi_grid_sync(bar + 2, ++phase)) goto fail;
      for (int stage = 0; stage < kBisectStages; ++stage) {
        if (!ld_ag(&sel->bisect)) break;  // the search ended without (more) bisection
        if (blk < nh) multi_tree_body(a, d, keys, state, sel, cum, blk, nh);
        if (!multi_grid_sync(bar + 2, ++phase)) goto fail;
        if (blk == 0)
          multi_bisect_body(a, d, state, sel, keys, cum, ext, nh, lower, upper,
                            stage + 1 == kBisectStages ? 1 : 0, 0, 1);
        if (!multi_grid_sync(bar + 2, ++phase)) goto fail;
      }
    }
  }
  a.e = 0;
  a.v = dates.v[0];
  a.k = basis_count(a.v, cross);
  if (blk < nb)
    multi_final_body(a, bv, cash, ev, num, coeffs, lower, upper, norm, Y, price_part, blk, nb);
  if (!multi_grid_sync(bar + 2, ++phase)) goto fail;
  if (blk == 0) finalize_col<true>(price_part, nb, 1, 0, 1.0 / static_cast<double>(a.P), price_out);
  return;
fail:
  if (threadIdx.x == 0) {
    bar[1] = 1u;
    if (blk == 0) *price_out = __builtin_nan("");
  }
}

// ---------------------------------------------------------------------------
// Host side.
// ---------------------------------------------------------------------------
int multi_blocks(int paths) {
  const int want = (paths + kMultiThreads - 1) / kMultiThreads;
  const int cap = device_cu_count();
  return want < 1 ? 1 : (want > cap ? cap : want);
}

// Blocks of the ladder and tree histograms: few, so the one-block search
// sums few partials per count (the pass itself moves 8 B per path).
int hist_blocks(int paths) {
  const int nb = multi_blocks(paths);
  return nb < kHistBlocks ? nb : kHistBlocks;
}

size_t al256(size_t x) { return (x + 255) & ~static_cast<size_t>(255); }

struct MultiLayout {
  size_t Y, mem, keys, cum, stats, gram, net, hist, ext, price, norm, state, sel, bar, total;
};

MultiLayout multi_layout(int paths) {
  MultiLayout l{};
  const size_t P = static_cast<size_t>(paths);
  const size_t nb = static_cast<size_t>(multi_blocks(paths));
  size_t o = 0;
  l.Y = o;
  o += al256(8 * P);
  l.mem = o;
  o += al256(P);
  l.keys = o;
  o += al256(8 * P);
  l.cum = o;
  o += al256(4 * static_cast<size_t>(kTreeNodes) * nb);
  l.stats = o;
  o += al256(8 * 2 * kMaxBasisVars * nb);
  l.gram = o;
  o += al256(8 * kMaxGram * kMaxGram * nb);
  l.net = o;
  o += al256(8 * 2 * nb);
  l.hist = o;
  o += al256(8 * kLadderSlots * nb);
  l.ext = o;
  o += al256(sizeof(KeyExtent) * nb);
  l.price = o;
  o += al256(8 * nb);
  l.norm = o;
  o += al256(8 * static_cast<size_t>(kMaxRates) * kMaxDepth * 2 * kMaxBasisVars);
  l.state = o;
  o += al256(sizeof(DepthState) * kMaxDepth);
  l.sel = o;
  o += al256(sizeof(SelectState));
  l.bar = o;  // error word, spare, then one barrier flag per block
  o += al256(4 * (2 + nb));
  l.total = o;
  return l;
}

}  // namespace
}  // namespace koo

using namespace koo;

extern "C" int ls_multi_workspace(int paths, size_t* bytes) {
  if (paths < 1) return set_error(KOO_ERR_INVALID, "ls_multi_workspace: paths must be >= 1");
  if (bytes) *bytes = multi_layout(paths).total;
  return KOO_OK;
}

extern "C" int ls_multi_first_pass_gpu(
    int n_ex, const float* basis_variables, const int* vars_each_step, int max_vars,
    const float* cash_flows, const float* exercise_values, const float* numeraires, int paths,
    int use_cross, int normalise, int max_regressions, int lower_path_cutoff, int selector,
    double lower_frac, double upper_frac, double sd_guess, double multiplier, void* workspace,
    double* coefficients, double* lower_cuts, double* upper_cuts, double* means, double* sds,
    double* price_out,
    void* stream) {
  const char* who = "ls_multi_first_pass_gpu";
  if (n_ex < 1 || n_ex > kMaxRates || paths < 1 || max_vars < 0 || max_vars > kMaxBasisVars ||
      !vars_each_step)
    return set_error(KOO_ERR_INVALID,
                     "%s: need 1<=exercise dates<=%d, paths>=1, 0<=max_vars<=%d (got %d, %d, %d)",
                     who, kMaxRates, kMaxBasisVars, n_ex, paths, max_vars);
  for (int e = 0; e < n_ex; ++e)
    if (vars_each_step[e] < 0 || vars_each_step[e] > max_vars)
      return set_error(KOO_ERR_INVALID, "%s: vars_each_step[%d] = %d outside [0, %d]", who, e,
                       vars_each_step[e], max_vars);
  const bool cross = use_cross != 0;
  const int maxd = basis_count(max_vars, cross);
  if (maxd + 1 > kMaxGram)
    return set_error(KOO_ERR_INVALID, "%s: %d basis functions exceed %d", who, maxd,
                     kMaxGram - 1);
  if (max_regressions < 1 || max_regressions > kMaxDepth)
    return set_error(KOO_ERR_INVALID, "%s: max_regressions %d outside [1, %d]", who,
                     max_regressions, kMaxDepth);
  if (lower_path_cutoff < 0)
    return set_error(KOO_ERR_INVALID, "%s: lower_path_cutoff %d < 0", who, lower_path_cutoff);
  if (selector != 0 && selector != 1)
    return set_error(KOO_ERR_INVALID, "%s: selector %d is neither 0 (fraction) nor 1 (sds)", who,
                     selector);
  if (!(sd_guess > 0.0))
    return set_error(KOO_ERR_INVALID, "%s: sd_guess %g must be > 0", who, sd_guess);
  if (selector == 0 && (!(multiplier > 0.0 && multiplier < 1.0) || !(lower_frac >= 0.0) ||
                        !(upper_frac >= lower_frac)))
    return set_error(KOO_ERR_INVALID,
                     "%s: the fraction selector needs 0 < multiplier < 1 and "
                     "0 <= lower_frac <= upper_frac (got %g, %g, %g)",
                     who, multiplier, lower_frac, upper_frac);
  if (!basis_variables || !cash_flows || !exercise_values || !numeraires || !workspace ||
      !coefficients || !lower_cuts || !upper_cuts || !price_out)
    return set_error(KOO_ERR_INVALID, "%s: null pointer", who);
  if (reinterpret_cast<uintptr_t>(workspace) % 256 != 0)
    return set_error(KOO_ERR_INVALID, "%s: workspace must be 256-byte aligned", who);
  const MultiLayout lay = multi_layout(paths);
  char* ws = static_cast<char*>(workspace);
  double* Y = reinterpret_cast<double*>(ws + lay.Y);
  unsigned char* mem = reinterpret_cast<unsigned char*>(ws + lay.mem);
  unsigned long long* keys = reinterpret_cast<unsigned long long*>(ws + lay.keys);
  int* cum = reinterpret_cast<int*>(ws + lay.cum);
  double* stats = reinterpret_cast<double*>(ws + lay.stats);
  double* gram = reinterpret_cast<double*>(ws + lay.gram);
  double* net = reinterpret_cast<double*>(ws + lay.net);
  double* hist = reinterpret_cast<double*>(ws + lay.hist);
  KeyExtent* ext = reinterpret_cast<KeyExtent*>(ws + lay.ext);
  double* price_part = reinterpret_cast<double*>(ws + lay.price);
  double* norm = reinterpret_cast<double*>(ws + lay.norm);
  DepthState* state = reinterpret_cast<DepthState*>(ws + lay.state);
  SelectState* sel = reinterpret_cast<SelectState*>(ws + lay.sel);
  const int R = max_regressions;
  const int nb = multi_blocks(paths);
  const int nh = hist_blocks(paths);
  hipStream_t st = as_stream(stream);
  MultiArgs a{};
  a.E = n_ex;
  a.P = paths;
  a.max_vars = max_vars;
  a.maxd = maxd;
  a.R = R;
  a.cross = cross ? 1 : 0;
  a.normalise = normalise ? 1 : 0;
  a.cutoff = lower_path_cutoff;
  a.selector = selector;
  a.lower_frac = lower_frac;
  a.upper_frac = upper_frac;
  a.sd_guess = sd_guess;
  a.multiplier = multiplier;

  // With KOO_LS_MULTI_PERSISTENT=1, one persistent launch (after zeroing
  // its barrier words) when all nb blocks can be resident: the same bodies
  // and blocks, bit for bit the same results.  Not the default: on the
  // config-3 path quantities it measures 5.0 ms per pass against 3.7 ms for
  // the launch per phase, the one-block band searches running 16 wavefronts
  // of the same serial search where their own launches run 4.
  {
    int per_cu = 0;
    const char* persistent = getenv("KOO_LS_MULTI_PERSISTENT");
    if (persistent && persistent[0] == '1' &&
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_multi_pass, kMultiThreads, 0) ==
            hipSuccess &&
        per_cu * device_cu_count() >= nb) {
      MultiDates dates{};
      for (int e = 0; e < n_ex; ++e) dates.v[e] = static_cast<signed char>(vars_each_step[e]);
      unsigned* bar = reinterpret_cast<unsigned*>(ws + lay.bar);
      hipLaunchKernelGGL(k_zero_words, dim3(1), dim3(256), 0, st, bar, 2 + nb);
      hipLaunchKernelGGL(k_multi_pass, dim3(nb), dim3(kMultiThreads), 0, st, a, dates, nb, nh,
                         basis_variables, cash_flows, exercise_values, numeraires, coefficients,
                         lower_cuts, upper_cuts, norm, means, sds, Y, mem, keys, cum, stats, gram,
                         net, hist, ext, price_part, state, sel, bar, price_out);
      return check_launch(who);
    }
  }

  // outputs start at zero (depths never reached stay zero)
  const size_t ncoef = static_cast<size_t>(n_ex) * R * maxd, ncut = static_cast<size_t>(n_ex) * R;
  const size_t nvar = static_cast<size_t>(n_ex) * R * max_vars;
  hipLaunchKernelGGL(k_fill_zero, dim3(grid_for(ncoef, 256, 256)), dim3(256), 0, st, coefficients,
                     ncoef);
  hipLaunchKernelGGL(k_fill_zero, dim3(grid_for(ncut, 256, 256)), dim3(256), 0, st, lower_cuts,
                     ncut);
  hipLaunchKernelGGL(k_fill_zero, dim3(grid_for(ncut, 256, 256)), dim3(256), 0, st, upper_cuts,
                     ncut);
  hipLaunchKernelGGL(k_fill_zero, dim3(grid_for(2 * nvar + 1, 256, 256)), dim3(256), 0, st, norm,
                     2 * nvar + 1);
  if (means && nvar)
    hipLaunchKernelGGL(k_fill_zero, dim3(grid_for(nvar, 256, 256)), dim3(256), 0, st, means, nvar);
  if (sds && nvar)
    hipLaunchKernelGGL(k_fill_zero, dim3(grid_for(nvar, 256, 256)), dim3(256), 0, st, sds, nvar);

  for (int e = n_ex - 1; e >= 0; --e) {
    a.e = e;
    a.v = vars_each_step[e];
    a.k = basis_count(a.v, cross);
    a.vn = e + 1 < n_ex ? vars_each_step[e + 1] : 0;
    a.kn = e + 1 < n_ex ? basis_count(a.vn, cross) : 0;
    for (int d = 0; d < R; ++d) {
      hipLaunchKernelGGL(k_multi_prep, dim3(nb), dim3(kMultiThreads), 0, st, a, d,
                         basis_variables, cash_flows, exercise_values, numeraires, coefficients,
                         lower_cuts, upper_cuts, norm, Y, mem, keys, state, stats);
      if (a.k + 1 <= 8)
        hipLaunchKernelGGL(k_multi_gram<2>, dim3(nb), dim3(kMultiThreads), 0, st, a, d,
                           basis_variables, cash_flows, Y, mem, state, stats, norm, means, sds,
                           gram);
      else
        hipLaunchKernelGGL(k_multi_gram<1>, dim3(nb), dim3(kMultiThreads), 0, st, a, d,
                           basis_variables, cash_flows, Y, mem, state, stats, norm, means, sds,
                           gram);
      hipLaunchKernelGGL(k_multi_net, dim3(nb), dim3(kMultiThreads), 0, st, a, d,
                         basis_variables, exercise_values, mem, state, norm, gram, coefficients,
                         keys, net);
      hipLaunchKernelGGL(k_multi_ladder, dim3(nh), dim3(kMultiThreads), 0, st, a, d, keys, net, nb,
                         state, hist, ext);
      hipLaunchKernelGGL(k_multi_select, dim3(1), dim3(kSearchThreads), 0, st, a, d, keys, net,
                         nb, hist, ext, nh, state, sel, lower_cuts, upper_cuts);
      for (int stage = 0; stage < kBisectStages; ++stage) {
        hipLaunchKernelGGL(k_multi_tree, dim3(nh), dim3(kMultiThreads), 0, st, a, d, keys, state,
                           sel, cum);
        hipLaunchKernelGGL(k_multi_bisect, dim3(1), dim3(kSearchThreads), 0, st, a, d, state, sel,
                           keys, cum, ext, nh, lower_cuts, upper_cuts,
                           stage + 1 == kBisectStages ? 1 : 0);
      }
    }
  }
  a.e = 0;
  a.v = vars_each_step[0];
  a.k = basis_count(a.v, cross);
  hipLaunchKernelGGL(k_multi_final, dim3(nb), dim3(kMultiThreads), 0, st, a, basis_variables,
                     cash_flows, exercise_values, numeraires, coefficients, lower_cuts,
                     upper_cuts, norm, Y, price_part);
  hipLaunchKernelGGL(k_finalize_cols, dim3(1), dim3(256), 0, st, price_part, nb, 1,
                     1.0 / static_cast<double>(paths), price_out);
  return check_launch(who);
}
