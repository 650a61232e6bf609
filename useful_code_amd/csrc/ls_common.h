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

// Sums the entry-major partials part[t * nb + b] over b < nb into sum[t]
// (LDS) for t < nn: each wavefront takes 4 entries at a time and each lane 4
// blocks of each (all 16 loads in flight, coalesced), then wavefront sums.
// Contains a barrier.
__device__ __forceinline__ void reduce_partials(const double* __restrict__ part, int nb, int nn,
                                                double* sum) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  for (int t0 = wave * 4; t0 < nn; t0 += nw * 4) {
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    for (int b0 = 0; b0 < nb; b0 += 256) {
      double v[4][4];
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int t = t0 + u, b = b0 + lane + 64 * r;
          v[u][r] = (t < nn && b < nb) ? part[static_cast<size_t>(t) * nb + b] : 0.0;
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

// The same solve with the factor in registers (K known at compile time, all
// loops unrolled): G in LDS, beta out.  Returns false (beta = 0) when a
// pivot is not positive.
template <int K>
__device__ bool cholesky_rhs_reg(const double* G, int ld, double inv, double* beta) {
  double L[K][K], y[K];
#pragma unroll
  for (int i = 0; i < K; ++i) {
#pragma unroll
    for (int j = 0; j <= i; ++j) {
      double s = G[i * ld + j] * inv;
#pragma unroll
      for (int q = 0; q < j; ++q) s -= L[i][q] * L[j][q];
      if (i == j) {
        if (!(s > 0.0)) {
#pragma unroll
          for (int r = 0; r < K; ++r) beta[r] = 0.0;
          return false;
        }
        L[i][i] = sqrt(s);
      } else {
        L[i][j] = s / L[j][j];
      }
    }
  }
#pragma unroll
  for (int i = 0; i < K; ++i) {
    double s = G[i * ld + K] * inv;
#pragma unroll
    for (int q = 0; q < i; ++q) s -= L[i][q] * y[q];
    y[i] = s / L[i][i];
  }
  double x[K];
#pragma unroll
  for (int i = K - 1; i >= 0; --i) {
    double s = y[i];
#pragma unroll
    for (int q = i + 1; q < K; ++q) s -= L[q][i] * x[q];
    x[i] = s / L[i][i];
  }
#pragma unroll
  for (int i = 0; i < K; ++i) beta[i] = x[i];
  return true;
}

// LS normal equations from a (k+1) x (k+1) Gram matrix whose row 0 is the
// constant function (so G[0][0] is the path count) and row k the target:
// mean-scaled as in GenerateLSRegressionMatrixSelectively_gold
// (gold/LS_regression.h:431-519); beta = 0 when no path entered or the
// system is not positive definite.
__device__ void solve_normal_equations(const double* G, int k, double* beta, double* L,
                                       double* y) {
  const double cnt = G[0];
  if (!(cnt > 0.0)) {
    for (int q = 0; q < k; ++q) beta[q] = 0.0;
    return;
  }
  const double inv = 1.0 / cnt;
  const int ld = k + 1;
  switch (k) {  // the basis sizes of 0..4 variables, cross or plain
    case 1: cholesky_rhs_reg<1>(G, ld, inv, beta); return;
    case 3: cholesky_rhs_reg<3>(G, ld, inv, beta); return;
    case 5: cholesky_rhs_reg<5>(G, ld, inv, beta); return;
    case 6: cholesky_rhs_reg<6>(G, ld, inv, beta); return;
    case 7: cholesky_rhs_reg<7>(G, ld, inv, beta); return;
    default: cholesky_rhs(G, ld, k, inv, beta, L, y); return;
  }
}

// Every block of a date kernel solves the later date's normal equations
// itself (same partials, same order, so the same coefficients): the sum of
// the entry-major partials into LDS, the FP64 solve by lane 0, the
// coefficients left in beta (LDS).  Block 0 also stores them.  Contains
// barriers.
__device__ __forceinline__ void solve_date(const double* __restrict__ part, int nb, int k,
                                           double* gram, double* chol, double* beta,
                                           double* __restrict__ coeffs_row, int maxd, int block) {
  reduce_partials(part, nb, (k + 1) * (k + 1), gram);
  if (threadIdx.x == 0) {
    solve_normal_equations(gram, k, beta, chol, chol + kMaxGram * kMaxGram);
    if (block == 0)
      for (int i = 0; i < maxd; ++i) coeffs_row[i] = i < k ? beta[i] : 0.0;
  }
  __syncthreads();
}

}  // namespace
}  // namespace koo
