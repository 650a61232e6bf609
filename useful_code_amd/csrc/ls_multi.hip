// ls_multi.hip -- the multi-regression Longstaff-Schwartz first pass
// (generateMultiRegressionCoefficientsViaLS_gpu_flexisolve, LS_main_cula.h:
// 565-838, with RegressionSelectorFraction, RegressionSelectorConcrete.cu:
// 102-233; the gold it is checked against is
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
//                   values of the members as 64-bit sort keys |net| (FP64
//                   bits; non-members get a sentinel above every key) and
//                   FP64 partial sums of net and net^2.
//   radix sort      rocPRIM radix_sort_keys over the keys.
//   k_multi_search  one block: the selector's band search (grow, shrink,
//                   bisect on a symmetric band) with exact counts read off
//                   the sorted keys (a 4096-entry sample index in LDS, then
//                   one segment of keys), the cuts, and the next depth's
//                   member count and activity.
// The whole pass computes in FP64 from the FP32 inputs, so the band search
// takes the same steps as the FP64 gold: counts are exact and the net
// values agree to rounding.  Depths past the last active one return at once;
// the host launches every depth, so no result is read back during the pass.
#include <rocprim/device/device_radix_sort.hpp>

#include "ls_common.h"

namespace koo {
namespace {

typedef double doublex4 __attribute__((ext_vector_type(4)));

constexpr int kMultiThreads = 1024;  // multi-block kernels (16 wavefronts)
constexpr int kSearchThreads = 256;
constexpr int kMaxDepth = 8;         // regressions per exercise date
constexpr int kIndexMax = 4096;      // sample index of the sorted keys (LDS)
constexpr int kSelectMaxSteps = 1000;
constexpr double kSvdTolerance = 1e-8;  // gold/MultiLS_regression.cpp:172
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
    beta[t] = coeffs[static_cast<size_t>(q) * a.R * a.maxd + t];
  for (int t = threadIdx.x; t < a.R * 2 * a.max_vars; t += blockDim.x)
    nrm[t] = norm[norm_at(q, a.R, 0, 0, a.max_vars) + t];
  for (int t = threadIdx.x; t < a.R; t += blockDim.x) {
    lo[t] = lower[static_cast<size_t>(q) * a.R + t];
    hi[t] = upper[static_cast<size_t>(q) * a.R + t];
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
    part[static_cast<size_t>(threadIdx.x) * nblocks + block] = s;
  }
}

// ---------------------------------------------------------------------------
// Membership, decision and variable sums.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kMultiThreads)
    k_multi_prep(const MultiArgs a, int depth, const float* __restrict__ bv,
                 const float* __restrict__ cash, const float* __restrict__ ev,
                 const float* __restrict__ num, const double* __restrict__ coeffs,
                 const double* __restrict__ lower, const double* __restrict__ upper,
                 const double* __restrict__ norm, double* __restrict__ Y,
                 unsigned char* __restrict__ mem, const unsigned long long* __restrict__ keys,
                 DepthState* __restrict__ state, double* __restrict__ stats_part) {
  __shared__ double beta[kMaxDepth * kMaxGram];
  __shared__ double nrm[kMaxDepth * 2 * kMaxBasisVars];
  __shared__ double lo[kMaxDepth], hi[kMaxDepth];
  __shared__ double red[16 * 8];
  const size_t P = static_cast<size_t>(a.P);
  unsigned long long cutbits = 0;
  if (depth == 0) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      state[0].active = 1;
      state[0].n_pre = a.P;
      state[0].cutbits = kKeySentinel;
    }
    if (a.e + 1 < a.E) load_decision(a, a.e + 1, coeffs, lower, upper, norm, beta, nrm, lo, hi);
  } else {
    if (!state[depth].active) return;  // uniform over the grid
    cutbits = state[depth].cutbits;
  }
  double sums[2 * kMaxBasisVars];
#pragma unroll
  for (int i = 0; i < 2 * kMaxBasisVars; ++i) sums[i] = 0.0;
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < a.P; p += gridDim.x * blockDim.x) {
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
      member = keys[p] <= cutbits;
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
    block_sums(sums, 2 * a.v, red, stats_part, blockIdx.x, gridDim.x);
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
    part[static_cast<size_t>(t) * nblocks + block] = s;
  }
}

template <int GROUPS>
__global__ void __launch_bounds__(kMultiThreads)
    k_multi_gram(const MultiArgs a, int depth, const float* __restrict__ bv,
                 const float* __restrict__ cash, const double* __restrict__ Y,
                 const unsigned char* __restrict__ mem, const DepthState* __restrict__ state,
                 const double* __restrict__ stats_part, double* __restrict__ norm,
                 double* __restrict__ means_out, double* __restrict__ sds_out,
                 double* __restrict__ gram_part) {
  __shared__ double lds_d[(kMultiThreads / 64) * 256];
  __shared__ double sums[2 * kMaxBasisVars];
  __shared__ double msd[2 * kMaxBasisVars];
  if (!state[depth].active) return;
  const double n_pre = static_cast<double>(state[depth].n_pre);
  if (a.normalise && a.v > 0) reduce_partials(stats_part, gridDim.x, 2 * a.v, sums);
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
    if (blockIdx.x == 0 && i < a.v) {
      norm[norm_at(a.e, a.R, depth, 0, a.max_vars) + i] = m;
      norm[norm_at(a.e, a.R, depth, 1, a.max_vars) + i] = sd;
      const size_t o = (static_cast<size_t>(a.e) * a.R + depth) * a.max_vars + i;
      if (means_out) means_out[o] = m;
      if (sds_out) sds_out[o] = sd;
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
  const int stride = gridDim.x * nw * 64;
  for (int base = (blockIdx.x * nw + wave) * 64; base < a.P; base += stride) {
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
  block_gram_f64<GROUPS>(acc, n, lds_d, gram_part, blockIdx.x, gridDim.x);
}

// ---------------------------------------------------------------------------
// Normal-equation solve and net continuation values.
// ---------------------------------------------------------------------------

// Eigen-solve of the symmetric k x k system M x = b with eigenvalues below
// tol times the largest dropped (the truncated SVD solve of
// SVDSquareMatrix::SVDSolve, gold/math/svd_gold.cpp:560-596, for a symmetric
// matrix): cyclic Jacobi (Golub and Van Loan, Matrix Computations, 4th ed.,
// 8.5.2) on A (LDS, ld 16, overwritten) with eigenvectors in V (LDS).
// Single lane.
__device__ void jacobi_truncated(int k, double* A, double* V, const double* b, double tol,
                                 double* x) {
  for (int i = 0; i < k; ++i)
    for (int j = 0; j < k; ++j) V[i * 16 + j] = i == j ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, tot = 0.0;
    for (int i = 0; i < k; ++i)
      for (int j = 0; j < k; ++j) {
        const double q = A[i * 16 + j] * A[i * 16 + j];
        tot += q;
        if (i != j) off += q;
      }
    if (off <= 1e-30 * tot) break;
    for (int p = 0; p < k - 1; ++p)
      for (int q = p + 1; q < k; ++q) {
        const double apq = A[p * 16 + q];
        if (apq == 0.0) continue;
        const double tau = (A[q * 16 + q] - A[p * 16 + p]) / (2.0 * apq);
        const double t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
        const double c = 1.0 / sqrt(1.0 + t * t), s = t * c;
        for (int r = 0; r < k; ++r) {
          const double arp = A[r * 16 + p], arq = A[r * 16 + q];
          A[r * 16 + p] = c * arp - s * arq;
          A[r * 16 + q] = s * arp + c * arq;
        }
        for (int r = 0; r < k; ++r) {
          const double apr = A[p * 16 + r], aqr = A[q * 16 + r];
          A[p * 16 + r] = c * apr - s * aqr;
          A[q * 16 + r] = s * apr + c * aqr;
        }
        for (int r = 0; r < k; ++r) {
          const double vrp = V[r * 16 + p], vrq = V[r * 16 + q];
          V[r * 16 + p] = c * vrp - s * vrq;
          V[r * 16 + q] = s * vrp + c * vrq;
        }
      }
  }
  double lmax = 0.0;
  for (int i = 0; i < k; ++i) lmax = fmax(lmax, fabs(A[i * 16 + i]));
  for (int i = 0; i < k; ++i) x[i] = 0.0;
  for (int i = 0; i < k; ++i) {
    const double l = A[i * 16 + i];
    if (!(lmax > 0.0) || fabs(l) < tol * lmax) continue;
    double e = 0.0;
    for (int r = 0; r < k; ++r) e += V[r * 16 + i] * b[r];
    e /= l;
    for (int r = 0; r < k; ++r) x[r] += e * V[r * 16 + i];
  }
}

// The normal equations M beta = t with M = G[0..k)[0..k) / n, t = G[.][k] / n
// (G: (k+1) x (k+1) Gram of basis and target).  Cholesky first: when it
// succeeds and 1 / trace(M^-1) >= tol * trace(M) -- a lower bound of the
// smallest eigenvalue against an upper bound of the largest -- no eigenvalue
// is below tol times the largest, the truncated solve keeps every direction
// and equals the Cholesky solution.  Otherwise the truncated Jacobi solve.
// Single lane; A, V are 16 x 16 LDS scratch, b and z 16.
__device__ void solve_multi(const double* G, int k, double n, double* beta, double* A, double* V,
                            double* b, double* z) {
  const int ld = k + 1;
  const double inv = 1.0 / n;
  for (int i = 0; i < k; ++i) {
    for (int j = 0; j < k; ++j) A[i * 16 + j] = G[i * ld + j] * inv;
    b[i] = G[i * ld + k] * inv;
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

__global__ void __launch_bounds__(kMultiThreads)
    k_multi_net(const MultiArgs a, int depth, const float* __restrict__ bv,
                const float* __restrict__ ev, const unsigned char* __restrict__ mem,
                const DepthState* __restrict__ state, const double* __restrict__ norm,
                const double* __restrict__ gram_part, double* __restrict__ coeffs,
                unsigned long long* __restrict__ keys, double* __restrict__ net_part) {
  __shared__ double gram[kMaxGram * kMaxGram];
  __shared__ double A[16 * 16], V[16 * 16], rhs[16], tmp[16];
  __shared__ double beta[kMaxGram];
  __shared__ double msd[2 * kMaxBasisVars];
  __shared__ double red[16 * 8];
  if (!state[depth].active) return;
  const int n = a.k + 1;
  reduce_partials(gram_part, gridDim.x, n * n, gram);
  if (threadIdx.x == 0) {
    solve_multi(gram, a.k, static_cast<double>(state[depth].n_pre), beta, A, V, rhs, tmp);
    if (blockIdx.x == 0)
      for (int i = 0; i < a.maxd; ++i)
        coeffs[(static_cast<size_t>(a.e) * a.R + depth) * a.maxd + i] = i < a.k ? beta[i] : 0.0;
  }
  if (threadIdx.x < 2 * kMaxBasisVars) {
    const int which = threadIdx.x / kMaxBasisVars, i = threadIdx.x % kMaxBasisVars;
    msd[threadIdx.x] = i < a.v ? norm[norm_at(a.e, a.R, depth, which, a.max_vars) + i]
                               : (which ? 1.0 : 0.0);
  }
  __syncthreads();
  const size_t P = static_cast<size_t>(a.P);
  double s[2] = {0.0, 0.0};
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < a.P; p += gridDim.x * blockDim.x) {
    unsigned long long key = kKeySentinel;
    if (mem[p]) {
      double x[kMaxBasisVars];
      norm_vars(bv, a.e, a.max_vars, P, p, a.v, msd, msd + kMaxBasisVars, x);
      const double net = basis_dot(a.v, a.k, a.cross != 0, x, beta) -
                         static_cast<double>(ev[static_cast<size_t>(a.e) * P + p]);
      s[0] += net;
      s[1] += net * net;
      key = static_cast<unsigned long long>(__double_as_longlong(fabs(net)));
    }
    keys[p] = key;
  }
  block_sums(s, 2, red, net_part, blockIdx.x, gridDim.x);
}

// ---------------------------------------------------------------------------
// The selector's band search on the sorted keys.
// ---------------------------------------------------------------------------

// Number of member keys <= gb (sorted keys, members first): the sample index
// idx[j] = keys[j * s] (j < m) locates the segment, one wave counts it.
// Called by the 64 lanes of one wave; uniform result.
__device__ __forceinline__ long long count_le(const unsigned long long* __restrict__ keys,
                                              long long n_pre, const unsigned long long* idx,
                                              int m, long long s, unsigned long long gb) {
  const int lane = threadIdx.x & 63;
  if (n_pre == 0) return 0;
  const int nch = (m + 63) / 64;
  const int last = lane * 64 + 63 < m ? lane * 64 + 63 : m - 1;
  const int c = __popcll(__ballot(lane < nch && idx[last] <= gb));
  int pos;
  if (c == nch) {
    pos = m;
  } else {
    const int j = c * 64 + lane;
    pos = c * 64 + __popcll(__ballot(j < m && idx[j] <= gb));
  }
  if (pos == 0) return 0;
  const long long start = static_cast<long long>(pos - 1) * s;
  const long long end = start + s < n_pre ? start + s : n_pre;
  long long cnt = start;
  for (long long i0 = start; i0 < end; i0 += 64) {
    const long long i = i0 + lane;
    cnt += __popcll(__ballot(i < end && keys[i] <= gb));
  }
  return cnt;
}

// Members with |net| < g.
__device__ __forceinline__ long long count_open(const unsigned long long* __restrict__ keys,
                                                long long n_pre, const unsigned long long* idx,
                                                int m, long long s, double g) {
  if (!(g > 0.0)) return 0;
  return count_le(keys, n_pre, idx, m, s,
                  static_cast<unsigned long long>(__double_as_longlong(g)) - 1ull);
}

// Members with |net| <= g (the gold's -g <= net <= g).
__device__ __forceinline__ long long count_band(const unsigned long long* __restrict__ keys,
                                                long long n_pre, const unsigned long long* idx,
                                                int m, long long s, double g) {
  if (!(g >= 0.0)) return 0;
  return count_le(keys, n_pre, idx, m, s,
                  static_cast<unsigned long long>(__double_as_longlong(g + 0.0)));
}

__global__ void __launch_bounds__(kSearchThreads)
    k_multi_search(const MultiArgs a, int depth, const unsigned long long* __restrict__ keys,
                   const double* __restrict__ net_part, int nb, DepthState* __restrict__ state,
                   double* __restrict__ lower, double* __restrict__ upper) {
  __shared__ unsigned long long idx[kIndexMax];
  __shared__ double sums[2];
  if (!state[depth].active) {
    if (threadIdx.x == 0 && depth + 1 < a.R) state[depth + 1].active = 0;
    return;
  }
  const long long n_pre = state[depth].n_pre;
  reduce_partials(net_part, nb, 2, sums);
  const long long s = n_pre > kIndexMax ? (n_pre + kIndexMax - 1) / kIndexMax : 1;
  const int m = static_cast<int>((n_pre + s - 1) / s);
  for (int j = threadIdx.x; j < m; j += blockDim.x) idx[j] = keys[j * s];
  __syncthreads();
  if (threadIdx.x >= 64) return;
  if (a.selector == 1) {
    // RegressionSelectorStandardDeviations (RegressionSelectorConcrete.cu:19-90):
    // the open band of sd_guess standard deviations
    const double mean = sums[0] / static_cast<double>(n_pre);
    const double sd = sqrt(sums[1] / static_cast<double>(n_pre) - mean * mean);
    const double cut = a.sd_guess * sd;
    const long long after = count_open(keys, n_pre, idx, m, s, cut);
    if (threadIdx.x == 0) {
      if (after > a.cutoff) {
        lower[static_cast<size_t>(a.e) * a.R + depth] = -cut;
        upper[static_cast<size_t>(a.e) * a.R + depth] = cut;
      }
      if (depth + 1 < a.R) {
        state[depth + 1].active = after > a.cutoff ? 1 : 0;
        state[depth + 1].n_pre = after;
        state[depth + 1].cutbits =
            cut > 0.0 ? static_cast<unsigned long long>(__double_as_longlong(cut)) - 1ull : 0ull;
      }
    }
    return;
  }
  // RegressionSelector_gold_Fraction::Select (Regression_Selector_concrete_gold.cpp:13-183)
  const long long lowerT = static_cast<long long>(static_cast<double>(n_pre) * a.lower_frac);
  const long long upperT = static_cast<long long>(static_cast<double>(n_pre) * a.upper_frac);
  double cut;
  if (n_pre > lowerT && n_pre < upperT) {
    cut = 10000.0;
  } else {
    const double mean = sums[0] / static_cast<double>(n_pre);
    const double sd = sqrt(sums[1] / static_cast<double>(n_pre) - mean * mean);
    double guess = sd * a.sd_guess;
    long long count = count_band(keys, n_pre, idx, m, s, guess);
    int it = 0;
    while (count < lowerT && it++ < kSelectMaxSteps) {
      guess /= a.multiplier;
      count = count_band(keys, n_pre, idx, m, s, guess);
    }
    if (count < upperT) {
      cut = guess;
    } else {
      double low = guess * a.multiplier;
      count = count_band(keys, n_pre, idx, m, s, low);
      it = 0;
      while (count > upperT && it++ < kSelectMaxSteps) {
        low *= a.multiplier;
        count = count_band(keys, n_pre, idx, m, s, low);
      }
      if (count > lowerT) {
        cut = low;
      } else {
        double high = guess;
        guess = 0.5 * (low + high);
        count = count_band(keys, n_pre, idx, m, s, guess);
        it = 0;
        while ((count < lowerT || count > upperT) && (high - low) > 1e-6 &&
               it++ < kSelectMaxSteps) {
          if (count < lowerT)
            low = guess;
          else
            high = guess;
          guess = 0.5 * (low + high);
          count = count_band(keys, n_pre, idx, m, s, guess);
        }
        cut = guess;
      }
    }
  }
  const long long after = count_band(keys, n_pre, idx, m, s, cut);
  if (threadIdx.x == 0) {
    if (after > a.cutoff) {
      lower[static_cast<size_t>(a.e) * a.R + depth] = -cut;
      upper[static_cast<size_t>(a.e) * a.R + depth] = cut;
    }
    if (depth + 1 < a.R) {
      state[depth + 1].active = after > a.cutoff ? 1 : 0;
      state[depth + 1].n_pre = after;
      state[depth + 1].cutbits =
          cut >= 0.0 ? static_cast<unsigned long long>(__double_as_longlong(cut + 0.0)) : 0ull;
    }
  }
}

// Date 0's decision and sum_p Y_p / B_0 into block partials.
__global__ void __launch_bounds__(kMultiThreads)
    k_multi_final(const MultiArgs a, const float* __restrict__ bv, const float* __restrict__ cash,
                  const float* __restrict__ ev, const float* __restrict__ num,
                  const double* __restrict__ coeffs, const double* __restrict__ lower,
                  const double* __restrict__ upper, const double* __restrict__ norm,
                  const double* __restrict__ Y, double* __restrict__ part) {
  __shared__ double beta[kMaxDepth * kMaxGram];
  __shared__ double nrm[kMaxDepth * 2 * kMaxBasisVars];
  __shared__ double lo[kMaxDepth], hi[kMaxDepth];
  __shared__ double red[16];
  load_decision(a, 0, coeffs, lower, upper, norm, beta, nrm, lo, hi);
  const size_t P = static_cast<size_t>(a.P);
  double s = 0.0;
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < a.P; p += gridDim.x * blockDim.x) {
    const double y = decide(a, 0, a.v, a.k, bv, cash, ev, P, p, Y[p], beta, nrm, lo, hi);
    s += y / static_cast<double>(num[p]);
  }
  block_sum_to(s, red, part + blockIdx.x);
}

__global__ void k_fill_zero(double* __restrict__ a, size_t n) {
  for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<size_t>(gridDim.x) * blockDim.x)
    a[i] = 0.0;
}

// ---------------------------------------------------------------------------
// Host side.
// ---------------------------------------------------------------------------
int multi_blocks(int paths) {
  const int want = (paths + kMultiThreads - 1) / kMultiThreads;
  const int cap = device_cu_count();
  return want < 1 ? 1 : (want > cap ? cap : want);
}

size_t al256(size_t x) { return (x + 255) & ~static_cast<size_t>(255); }

size_t sort_temp_bytes(int paths) {
  size_t bytes = 0;
  (void)rocprim::radix_sort_keys(nullptr, bytes, static_cast<const unsigned long long*>(nullptr),
                           static_cast<unsigned long long*>(nullptr),
                           static_cast<unsigned int>(paths), 0, 63);
  return bytes;
}

struct MultiLayout {
  size_t Y, mem, keys, sorted, stats, gram, net, price, norm, state, sort_tmp, sort_bytes, total;
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
  l.sorted = o;
  o += al256(8 * P);
  l.stats = o;
  o += al256(8 * 2 * kMaxBasisVars * nb);
  l.gram = o;
  o += al256(8 * kMaxGram * kMaxGram * nb);
  l.net = o;
  o += al256(8 * 2 * nb);
  l.price = o;
  o += al256(8 * nb);
  l.norm = o;
  o += al256(8 * static_cast<size_t>(kMaxRates) * kMaxDepth * 2 * kMaxBasisVars);
  l.state = o;
  o += al256(sizeof(DepthState) * kMaxDepth);
  l.sort_tmp = o;
  l.sort_bytes = sort_temp_bytes(paths);
  o += al256(l.sort_bytes);
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
  unsigned long long* sorted = reinterpret_cast<unsigned long long*>(ws + lay.sorted);
  double* stats = reinterpret_cast<double*>(ws + lay.stats);
  double* gram = reinterpret_cast<double*>(ws + lay.gram);
  double* net = reinterpret_cast<double*>(ws + lay.net);
  double* price_part = reinterpret_cast<double*>(ws + lay.price);
  double* norm = reinterpret_cast<double*>(ws + lay.norm);
  DepthState* state = reinterpret_cast<DepthState*>(ws + lay.state);
  void* sort_tmp = ws + lay.sort_tmp;
  const int R = max_regressions;
  const int nb = multi_blocks(paths);
  hipStream_t st = as_stream(stream);

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
      size_t sort_bytes = lay.sort_bytes;
      const hipError_t se = rocprim::radix_sort_keys(sort_tmp, sort_bytes, keys, sorted,
                                                     static_cast<unsigned int>(paths), 0, 63, st);
      if (se != hipSuccess)
        return set_error(KOO_ERR_HIP, "%s: radix sort failed: %s", who, hipGetErrorString(se));
      hipLaunchKernelGGL(k_multi_search, dim3(1), dim3(kSearchThreads), 0, st, a, d, sorted, net,
                         nb, state, lower_cuts, upper_cuts);
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
