// ls_common.h -- device pieces shared by the LS kernels (ls.hip, ls_multi.hip):
// the quadratic basis, the MFMA Gram contraction into FP64 block partials,
// the FP64 reduction of those partials and the FP64 Cholesky solve of the
// normal equations.
//
// MFMA layout used (v_mfma_f32_16x16x4f32, one 16x16x4 block per wave): lane l
// supplies A[l%16][l/16] and B[l/16][l%16]; result element i of lane l is
// D[4*(l/16)+i][l%16].  For a Gram matrix A == B: row m of A holds basis
// function m of the four paths k = 0..3.  With at most 8 rows, the 16 rows
// hold two groups of 4 paths (rows 0-7 and 8-15), so each instruction takes
// 8 paths and the two diagonal 8x8 blocks of D are added at the end.
#pragma once

#include "lmm_step.h"

namespace koo {
namespace {

using namespace lmm;

typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int kMaxBasisVars = 4;  // cross basis with 4 variables: 13 functions
constexpr int kMaxGram = 16;      // rows of one Gram matrix (basis functions + target)
constexpr int kLsThreads = 256;     // lane-per-path kernels
constexpr int kGramThreads = 1024;  // MFMA contraction kernels (16 wavefronts)

__device__ __forceinline__ floatx4 mfma16x16x4(float a, float b, floatx4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

template <class T>
__device__ __forceinline__ T pick4(int i, T x0, T x1, T x2, T x3) {
  return i == 0 ? x0 : (i == 1 ? x1 : (i == 2 ? x2 : x3));
}

// Quadratic basis function j of v variables.  Cross variant
// (quadraticPolynomialCrossDevice, LS_gpu.h:138-229): 1, x_i, then for v = 1
// x^2, for v = 2 x^2, xy, y^2, for v >= 3 x_i x_{(i+1) mod v} then x_i^2.
// Plain variant (LS_gpu.h:92-126): 1, x_0, x_0^2, x_1, x_1^2, ...
// Branch-free in j so lanes with different j stay converged.  T is float
// or double.
template <class T>
__device__ __forceinline__ T basis_value(int j, int v, bool cross, T x0, T x1, T x2, T x3) {
  if (j == 0) return T(1);
  if (!cross) {
    const T x = pick4((j - 1) >> 1, x0, x1, x2, x3);
    return (j & 1) ? x : x * x;
  }
  if (j <= v) return pick4(j - 1, x0, x1, x2, x3);
  if (v == 1) return x0 * x0;
  if (v == 2) return j == 3 ? x0 * x0 : (j == 4 ? x0 * x1 : x1 * x1);
  const int i = j - 1 - v;
  if (i < v) {
    const int i2 = (i + 1 == v) ? 0 : i + 1;
    return pick4(i, x0, x1, x2, x3) * pick4(i2, x0, x1, x2, x3);
  }
  const T x = pick4(i - v, x0, x1, x2, x3);
  return x * x;
}

// Basis function j of v variables (as basis_value; row k is the target and
// rows past it are zero) as the product s[ia] * s[ib] of a staged path
// vector s = [w, w x0, w x1, w x2, w x3, w t, 0, 1] (w the path's 0/1
// weight, t its target): since w * w == w, s[ia] * s[ib] is w times the
// function.  Lets a lane whose MFMA row is j read its two factors for any
// path without branching on j.
__device__ __forceinline__ void basis_factors(int j, int v, int k, bool cross, int& ia, int& ib) {
  constexpr int kOne = 7, kZero = 6, kTarget = 5;
  ia = 0;
  ib = kOne;
  if (j > k) {
    ia = ib = kZero;
  } else if (j == k) {
    ia = kTarget;
  } else if (j == 0) {
  } else if (!cross) {
    ia = ((j - 1) >> 1) + 1;
    if (!(j & 1)) ib = ia;
  } else if (j <= v) {
    ia = j;
  } else if (v == 1) {
    ia = ib = 1;
  } else if (v == 2) {
    ia = j == 5 ? 2 : 1;
    ib = j == 3 ? 1 : 2;
  } else {
    const int i = j - 1 - v;
    if (i < v) {
      ia = i + 1;
      ib = (i + 1 == v ? 0 : i + 1) + 1;
    } else {
      ia = ib = i - v + 1;
    }
  }
}

__host__ __device__ inline int basis_count(int v, bool cross) {
  if (!cross) return 1 + 2 * v;
  return v == 0 ? 1 : (v == 1 ? 3 : (v == 2 ? 6 : 1 + 3 * v));
}

// Adds the two diagonal 8x8 blocks (groups == 2) or takes the 16x16 result
// (groups == 1) of each wave's accumulator and sums the waves of the block in
// FP64 into the block's column of the entry-major partials:
// part[(i*n + j) * nblocks + block] for i, j < n.  Every lane of the block
// calls it.
template <int GROUPS>
__device__ __forceinline__ void block_gram(const floatx4& acc, int n, float* lds_d /* [nw][256] */,
                                           double* part, int block, int nblocks) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  float* d = lds_d + wave * 256;
#pragma unroll
  for (int i = 0; i < 4; ++i) d[(4 * (lane >> 4) + i) * 16 + (lane & 15)] = acc[i];
  __syncthreads();
  // 16 lanes per entry (4 entries per wavefront), lane q of the group adds
  // the tiles of waves q, q+16, ...; then a 16-lane butterfly.
  const int q = lane & 15;
  for (int t = wave * 4 + (lane >> 4); t - (lane >> 4) < n * n; t += nw * 4) {
    double s = 0.0;
    if (t < n * n) {
      const int i = t / n, j = t - i * n;
      for (int w = q; w < nw; w += 16) {
        const float* dw = lds_d + w * 256;
        s += static_cast<double>(dw[i * 16 + j]);
        if (GROUPS == 2) s += static_cast<double>(dw[(8 + i) * 16 + 8 + j]);
      }
    }
#pragma unroll
    for (int o = 8; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (q == 0 && t < n * n) part[static_cast<size_t>(t) * nblocks + block] = s;
  }
}

// Entries of the upper triangle (i <= j, row-major) of an n x n symmetric
// matrix, and the (i, j) of entry t.
__host__ __device__ constexpr int tri_count(int n) { return n * (n + 1) / 2; }
constexpr int kTriMax = tri_count(kMaxGram);
__device__ __forceinline__ void tri_index(int t, int n, int& i, int& j) {
  i = 0;
  while (t >= n - i) {
    t -= n - i;
    ++i;
  }
  j = i + t;
}

// Block partials exchanged between the blocks of one running kernel carry no
// counter: the buffer starts filled with this bit pattern (a signalling NaN,
// which no sum of finite terms and no arithmetic NaN has), a block's partial
// is there once the word no longer holds it.  64-bit aligned stores are
// single-copy atomic, so a reader sees either the pattern or the value.
constexpr unsigned long long kUnsetBits = 0x7FF4C0FFEE15C0DEull;
__device__ __forceinline__ bool is_unset(double v) {
  return static_cast<unsigned long long>(__double_as_longlong(v)) == kUnsetBits;
}

// Fills n 64-bit words with v (a kernel, not a memset, so a launch sequence
// that uses it can be captured in a graph).
__global__ void k_fill_u64(unsigned long long* __restrict__ p, size_t n, unsigned long long v) {
  for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<size_t>(gridDim.x) * blockDim.x)
    p[i] = v;
}

// The Gram of each wave's MFMA accumulators -- FP64 per lane (the LS first
// pass adds each MFMA's 4-path FP32 result into FP64, so its Gram carries no
// FP32 rounding beyond that of the products and 4-term sums) -- summed over
// the block's waves in FP64, the upper triangle of the n x n matrix into the
// block's column of the entry-major partials out[t * nblocks + block],
// t < tri_count(n).  GROUPS == 2 adds the two diagonal 8x8 blocks of D,
// GROUPS == 1 takes the 16x16 result.  Agent: agent-coherent stores (the
// persistent kernels).  Every lane of the block calls it; the caller
// synchronises before lds_d is written again.
template <int GROUPS, bool Agent = false>
__device__ __forceinline__ void block_gram_tri(const double (&acc)[4], int n,
                                               double* lds_d /* [nw][256] */, double* out,
                                               int block, int nblocks) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  double* d = lds_d + wave * 256;
#pragma unroll
  for (int i = 0; i < 4; ++i) d[(4 * (lane >> 4) + i) * 16 + (lane & 15)] = acc[i];
  __syncthreads();
  // 16 lanes per entry (4 entries per wavefront), lane q of the group adds
  // the tiles of waves q, q+16, ...; then a 16-lane butterfly.
  const int q = lane & 15;
  const int T = tri_count(n);
  for (int t = wave * 4 + (lane >> 4); t - (lane >> 4) < T; t += nw * 4) {
    double s = 0.0;
    if (t < T) {
      int i, j;
      tri_index(t, n, i, j);
      for (int w = q; w < nw; w += 16) {
        const double* dw = lds_d + w * 256;
        s += dw[i * 16 + j];
        if (GROUPS == 2) s += dw[(8 + i) * 16 + 8 + j];
      }
    }
#pragma unroll
    for (int o = 8; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (q == 0 && t < T) {
      if (Agent)
        store_agent(out + static_cast<size_t>(t) * nblocks + block, s);
      else
        out[static_cast<size_t>(t) * nblocks + block] = s;
    }
  }
}

// Sums the entry-major partials X[t * nb + b] over b < nb into sum[t] (LDS)
// for t < T, in a fixed order (each wavefront 4 entries at a time, each lane 4
// blocks of each, 16 loads in flight, then a wavefront butterfly), so every
// block that sums the same partials gets the same bits.  Poll: agent loads,
// re-read until no word holds kUnsetBits, at most `reads` times (1: a
// single try); returns false on every lane if a word was still unset.
// Contains barriers; every lane of the block calls it.
constexpr unsigned kBarrierPolls = 1u << 22;

template <bool Poll>
__device__ __forceinline__ bool sum_tri(const double* X, int nb, int T, double* sum,
                                        int* flag_lds, unsigned reads = kBarrierPolls) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  if (Poll) {
    if (threadIdx.x == 0) *flag_lds = 1;
    __syncthreads();
  }
  bool ok = true;
  for (int t0 = wave * 4; ok && t0 < T; t0 += nw * 4) {
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    for (int b0 = 0; ok && b0 < nb; b0 += 256) {
      double v[4][4];
      unsigned polls = 0;
      for (;;) {
        bool mine = true;
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int t = t0 + u, b = b0 + lane + 64 * r;
            const double* src = X + static_cast<size_t>(t) * nb + b;
            const bool in = t < T && b < nb;
            v[u][r] = in ? (Poll ? load_agent(src) : *src) : 0.0;
            if (Poll && in && is_unset(v[u][r])) mine = false;
          }
        if (!Poll || __all(mine)) break;
        if (++polls >= reads) {
          ok = false;
          break;
        }
        __builtin_amdgcn_s_sleep(2);
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) s[u] += (v[u][0] + v[u][1]) + (v[u][2] + v[u][3]);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const double tot = wave_sum_f64(s[u]);
      if (lane == 0 && t0 + u < T) sum[t0 + u] = tot;
    }
  }
  if (Poll && !ok && lane == 0) *flag_lds = 0;
  __syncthreads();
  if (!Poll) return true;
  ok = *flag_lds != 0;
  __syncthreads();
  return ok;
}

// Writes kUnsetBits over the block's own entries t < T of X (agent stores):
// the owner re-arms its words once every block has read them.
__device__ __forceinline__ void rearm_tri(double* X, int nb, int T, int block) {
  for (int t = threadIdx.x; t < T; t += blockDim.x)
    store_agent(X + static_cast<size_t>(t) * nb + block, __longlong_as_double(kUnsetBits));
}

// Sums the entry-major partials part[t * nb + b] over b < nb into sum[t]
// (LDS) for t < nn, on the wavefronts [w0, w0 + nws) of the block (the others
// return at once): each wavefront takes 4 entries at a time and each lane 4
// blocks of each (all 16 loads in flight, coalesced), then wavefront sums.
// An entry's sum does not depend on which wavefront forms it.  No barrier.
// Agent: agent-coherent loads.
template <bool Agent = false>
__device__ __forceinline__ void reduce_partials_waves(const double* __restrict__ part, int nb,
                                                      int nn, double* sum, int w0, int nws) {
  const int lane = threadIdx.x & 63, wave = (threadIdx.x >> 6) - w0;
  if (wave < 0 || wave >= nws) return;
  for (int t0 = wave * 4; t0 < nn; t0 += nws * 4) {
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    for (int b0 = 0; b0 < nb; b0 += 256) {
      double v[4][4];
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int t = t0 + u, b = b0 + lane + 64 * r;
          const double* src = part + static_cast<size_t>(t) * nb + b;
          v[u][r] = (t < nn && b < nb) ? (Agent ? load_agent(src) : *src) : 0.0;
        }
#pragma unroll
      for (int u = 0; u < 4; ++u) s[u] += (v[u][0] + v[u][1]) + (v[u][2] + v[u][3]);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const double tot = wave_sum_f64(s[u]);
      if (lane == 0 && t0 + u < nn) sum[t0 + u] = tot;
    }
  }
}

// reduce_partials_waves on every wavefront of the block, then a barrier.
template <bool Agent = false>
__device__ __forceinline__ void reduce_partials(const double* __restrict__ part, int nb, int nn,
                                                double* sum) {
  reduce_partials_waves<Agent>(part, nb, nn, sum, 0, blockDim.x >> 6);
  __syncthreads();
}

// Cholesky solve (FP64) of the k x k system in the leading block of G
// (leading dimension ld) with right-hand side in column k, both multiplied by
// `inv`.  Returns false, with beta = 0, when a pivot is not positive.
// Single lane; L (kMaxGram^2) and y (kMaxGram) are caller scratch (LDS), so
// the kernels need no private scratch memory.
__device__ bool cholesky_rhs(const double* G, int ld, int k, double inv, double* beta, double* L,
                             double* y) {
  for (int i = 0; i < k; ++i) {
    for (int j = 0; j <= i; ++j) {
      double s = G[i * ld + j] * inv;
      for (int q = 0; q < j; ++q) s -= L[i * kMaxGram + q] * L[j * kMaxGram + q];
      if (i == j) {
        if (!(s > 0.0)) {
          for (int q = 0; q < k; ++q) beta[q] = 0.0;
          return false;
        }
        L[i * kMaxGram + i] = sqrt(s);
      } else {
        L[i * kMaxGram + j] = s / L[j * kMaxGram + j];
      }
    }
  }
  for (int i = 0; i < k; ++i) {
    double s = G[i * ld + k] * inv;
    for (int q = 0; q < i; ++q) s -= L[i * kMaxGram + q] * y[q];
    y[i] = s / L[i * kMaxGram + i];
  }
  for (int i = k - 1; i >= 0; --i) {
    double s = y[i];
    for (int q = i + 1; q < k; ++q) s -= L[q * kMaxGram + i] * beta[q];
    beta[i] = s / L[i * kMaxGram + i];
  }
  return true;
}

// The gold's SVD tolerance for the LS normal equations: eigen-directions of
// M below this fraction of the largest eigenvalue are dropped
// (SVDSquareMatrix::SVDSolve, gold/math/svd_gold.cpp:560-596, called with
// 1e-8 by gold/MultiLS_regression.cpp:172 and by the reference GPU's host
// solver, LinearSolverConcrete_main.cu:127-188).
constexpr double kSvdTolerance = 1e-8;

// Eigen-solve of the symmetric k x k system M x = b with eigenvalues below
// tol times the largest dropped (the truncated SVD solve of a symmetric
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

// Cholesky solve with the factor in registers (K known at compile time, all
// loops unrolled) of M beta = t, M = G[0..K)[0..K) * inv, t = G[.][K] * inv.
// Succeeds only when M is positive definite and 1 / trace(M^-1) >= tol *
// trace(M) -- a lower bound of the smallest eigenvalue against an upper bound
// of the largest, so no eigenvalue is below tol times the largest and the
// truncated solve would keep every direction: then it equals the truncated
// solve.  Returns false (beta untouched) otherwise.
template <int K>
__device__ bool cholesky_rhs_reg(const double* G, int ld, double inv, double tol, double* beta) {
  double L[K][K], y[K], d[K];
  double trM = 0.0;
#pragma unroll
  for (int i = 0; i < K; ++i) {
#pragma unroll
    for (int j = 0; j <= i; ++j) {
      double s = G[i * ld + j] * inv;
      if (i == j) trM += s;
#pragma unroll
      for (int q = 0; q < j; ++q) s -= L[i][q] * L[j][q];
      if (i == j) {
        if (!(s > 0.0)) return false;
        L[i][i] = sqrt(s);
        d[i] = 1.0 / L[i][i];
      } else {
        L[i][j] = s * d[j];
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
      z[i] = s * d[i];
      trInv += z[i] * z[i];
    }
  }
  if (trInv * tol * trM > 1.0) return false;
#pragma unroll
  for (int i = 0; i < K; ++i) {
    double s = G[i * ld + K] * inv;
#pragma unroll
    for (int q = 0; q < i; ++q) s -= L[i][q] * y[q];
    y[i] = s * d[i];
  }
  double x[K];
#pragma unroll
  for (int i = K - 1; i >= 0; --i) {
    double s = y[i];
#pragma unroll
    for (int q = i + 1; q < K; ++q) s -= L[q][i] * x[q];
    x[i] = s * d[i];
  }
#pragma unroll
  for (int i = 0; i < K; ++i) beta[i] = x[i];
  return true;
}

// The same truncated eigen-solve with the matrix in registers (K known at
// compile time, loops over the rotation pairs unrolled; only the upper
// triangle is kept): each rotation J of the cyclic sweep (Golub and Van Loan
// 8.4.2: the (p,q) entry zeroed, the diagonal updated by -/+ t a_pq) is also
// applied to the right-hand side, b' = J^T b', and its (c, s) logged in LDS;
// at convergence the kept eigenvalues give y = Lambda^+ b' and the solution
// x = J_1 ... J_m y replays the log backwards.  M = G[0..K)[0..K) * inv,
// rhs G[.][K] * inv.  Single lane; rot: 2 * kJacobiSweeps * K(K-1)/2 LDS
// doubles.
constexpr int kJacobiSweeps = 32;

template <int K>
__device__ void jacobi_truncated_reg(const double* G, int ld, double inv, double tol, double* x,
                                     double* rot) {
  constexpr int NP = K * (K - 1) / 2;
  double a[K][K], b[K];
#pragma unroll
  for (int i = 0; i < K; ++i) {
#pragma unroll
    for (int j = i; j < K; ++j) a[i][j] = G[i * ld + j] * inv;
    b[i] = G[i * ld + K] * inv;
  }
  int sweeps = 0;
  for (; sweeps < kJacobiSweeps; ++sweeps) {
    double off = 0.0, tot = 0.0;
#pragma unroll
    for (int i = 0; i < K; ++i) {
      tot += a[i][i] * a[i][i];
#pragma unroll
      for (int j = i + 1; j < K; ++j) off += 2.0 * a[i][j] * a[i][j];
    }
    if (off <= 1e-30 * (tot + off)) break;
    double* log = rot + 2 * sweeps * NP;
    int idx = 0;
#pragma unroll
    for (int p = 0; p < K - 1; ++p) {
#pragma unroll
      for (int q = p + 1; q < K; ++q, ++idx) {
        const double apq = a[p][q];
        double c = 1.0, sn = 0.0;
        if (apq != 0.0) {
          const double tau = (a[q][q] - a[p][p]) / (2.0 * apq);
          const double t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
          c = 1.0 / sqrt(1.0 + t * t);
          sn = t * c;
          a[p][p] -= t * apq;
          a[q][q] += t * apq;
          a[p][q] = 0.0;
#pragma unroll
          for (int r = 0; r < K; ++r) {
            if (r == p || r == q) continue;
            double& arp = r < p ? a[r][p] : a[p][r];
            double& arq = r < q ? a[r][q] : a[q][r];
            const double vp = arp, vq = arq;
            arp = c * vp - sn * vq;
            arq = sn * vp + c * vq;
          }
          const double bp = b[p], bq = b[q];
          b[p] = c * bp - sn * bq;
          b[q] = sn * bp + c * bq;
        }
        log[2 * idx] = c;
        log[2 * idx + 1] = sn;
      }
    }
  }
  double lmax = 0.0;
#pragma unroll
  for (int i = 0; i < K; ++i) lmax = fmax(lmax, fabs(a[i][i]));
  double y[K];
#pragma unroll
  for (int i = 0; i < K; ++i)
    y[i] = (lmax > 0.0 && !(fabs(a[i][i]) < tol * lmax)) ? b[i] / a[i][i] : 0.0;
  for (int sw = sweeps - 1; sw >= 0; --sw) {
    const double* log = rot + 2 * sw * NP;
    int idx = NP - 1;
#pragma unroll
    for (int p = K - 2; p >= 0; --p) {
#pragma unroll
      for (int q = K - 1; q > p; --q, --idx) {
        const double c = log[2 * idx], sn = log[2 * idx + 1];
        const double yp = y[p], yq = y[q];
        y[p] = c * yp + sn * yq;
        y[q] = -sn * yp + c * yq;
      }
    }
  }
#pragma unroll
  for (int i = 0; i < K; ++i) x[i] = y[i];
}

// Scratch (LDS doubles) the truncated solves need: A and V (16 x 16) and b
// for the LDS form, the rotation log of the register form.
constexpr int kSolveScratch =
    (2 * 256 + 16) > (2 * kJacobiSweeps * 21) ? (2 * 256 + 16) : (2 * kJacobiSweeps * 21);

// LS normal equations from a (k+1) x (k+1) Gram matrix whose row 0 is the
// constant function (so G[0][0] is the path count) and row k the target,
// mean-scaled as in GenerateLSRegressionMatrixSelectively_gold
// (gold/LS_regression.h:431-519), solved like the reference's truncated SVD
// with tolerance 1e-8: the register Cholesky when the trace bound shows no
// direction is dropped (the usual case), else the Jacobi eigen-solve with the
// cut.  beta = 0 when no path entered.  Single lane; scratch: kSolveScratch
// LDS doubles.
__device__ void solve_normal_equations(const double* G, int k, double* beta, double* scratch) {
  const double cnt = G[0];
  if (!(cnt > 0.0)) {
    for (int q = 0; q < k; ++q) beta[q] = 0.0;
    return;
  }
  const double inv = 1.0 / cnt;
  const int ld = k + 1;
  bool done = false;
  switch (k) {  // the basis sizes of 0..3 variables, cross or plain
    case 1: done = cholesky_rhs_reg<1>(G, ld, inv, kSvdTolerance, beta); break;
    case 3: done = cholesky_rhs_reg<3>(G, ld, inv, kSvdTolerance, beta); break;
    case 5: done = cholesky_rhs_reg<5>(G, ld, inv, kSvdTolerance, beta); break;
    case 6: done = cholesky_rhs_reg<6>(G, ld, inv, kSvdTolerance, beta); break;
    case 7: done = cholesky_rhs_reg<7>(G, ld, inv, kSvdTolerance, beta); break;
    default: break;
  }
  if (done) return;
  switch (k) {
    case 1: jacobi_truncated_reg<1>(G, ld, inv, kSvdTolerance, beta, scratch); return;
    case 3: jacobi_truncated_reg<3>(G, ld, inv, kSvdTolerance, beta, scratch); return;
    case 5: jacobi_truncated_reg<5>(G, ld, inv, kSvdTolerance, beta, scratch); return;
    case 6: jacobi_truncated_reg<6>(G, ld, inv, kSvdTolerance, beta, scratch); return;
    case 7: jacobi_truncated_reg<7>(G, ld, inv, kSvdTolerance, beta, scratch); return;
    default: break;
  }
  double* A = scratch;
  double* V = scratch + 256;
  double* b = scratch + 512;
  for (int i = 0; i < k; ++i) {
    for (int j = 0; j < k; ++j) A[i * 16 + j] = G[i * ld + j] * inv;
    b[i] = G[i * ld + k] * inv;
  }
  jacobi_truncated(k, A, V, b, kSvdTolerance, beta);
}

// The normal equations of one date from the summed upper triangle (LDS,
// tri_count(k + 1) entries, the target last): the full (k+1) x (k+1) Gram in
// LDS, then the FP64 solve by lane 0 into beta (LDS).  Contains barriers;
// every lane of the block calls it.
__device__ __forceinline__ void solve_tri(const double* tri, int k, double* gram, double* scratch,
                                          double* beta) {
  const int n = k + 1;
  for (int t = threadIdx.x; t < tri_count(n); t += blockDim.x) {
    int i, j;
    tri_index(t, n, i, j);
    gram[i * n + j] = tri[t];
    gram[j * n + i] = tri[t];
  }
  __syncthreads();
  if (threadIdx.x == 0) solve_normal_equations(gram, k, beta, scratch);
  __syncthreads();
}

// Zeroes n 32-bit words (the barrier counters before a persistent launch;
// a kernel, not a memset, so the launch sequence can be captured in a graph).
__global__ void k_zero_words(unsigned* __restrict__ p, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0u;
}

}  // namespace
}  // namespace koo
