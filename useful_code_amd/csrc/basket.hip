// basket.hip -- multi-asset Black-Scholes paths on gfx950.
//
// * The reference's three launchers: correlation of factor variates into
//   asset log increments (correlated_paths_multistep_GPU), cumulative
//   log-Euler paths (multi_dim_BS_path_generator_kernel) and basket
//   weighting (basketWeightings_gpu).
// * k_basket_asian: the fused arithmetic-average basket Asian call with its
//   pathwise adjoint deltas and vegas.  A wavefront owns 64 paths; per step
//   the correlation z -> sigma sqrt(dt) L z of its 64 paths x A assets is a
//   small GEMM done on MFMA (v_mfma_f32_16x16x4f32, the triangular factor's
//   zero blocks skipped), and the log-Euler update, exponential and the
//   adjoint accumulations run on the result where the MFMA leaves it.
//
// Operand layout of one MFMA (M = 16 paths, N = 16 assets, K = 4 factors):
// lane l supplies A[i = l%16][k = l/16] and B[k = l/16][j = l%16]; result
// element r of lane l is D[4*(l/16) + r][l%16].  Path i of "path slice" q is
// wave path 4*i + q, so the four slices' A operands of one factor row are one
// 16-byte load per lane (wave paths 4*(l%16) .. 4*(l%16)+3).  Accordingly the
// D element (slice q, asset tile t, r) of lane l is wave path
// 16*(l/16) + 4*r + q and asset 16*t + l%16: the 16 lanes of a DPP row share
// the path and hold its assets, so the basket sum is a row reduction.
#include "lmm_step.h"

namespace koo {
namespace {

using namespace lmm;

typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int kBasketThreads = 256;
constexpr int kMaxBasketAssets = 64;  // four 16-asset tiles

__device__ __forceinline__ floatx4 mfma4(float a, float b, floatx4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// Sum over the 16 lanes of this lane's DPP row (every lane gets the total).
__device__ __forceinline__ float row16_sum(float v) {
  v += __shfl_xor(v, 8, 16);
  v += __shfl_xor(v, 4, 16);
  v += __shfl_xor(v, 2, 16);
  v += __shfl_xor(v, 1, 16);
  return v;
}

struct BasketArgs {
  int P, A, S, F;       // paths (multiple of 4), assets, steps, factor rows of z
  float dt, strike, disc, inv_steps;
};

// Per-asset constants, one set per asset tile for this lane's asset.  The
// path state is y = log(S/S(0)) (it starts at 0, so its FP32 rounding is
// that of a number of order sigma sqrt(t), not of log S ~ 4.6: accumulating
// log S itself costs ~1e-5 relative in the at-the-money payoff and ~2e-4 in
// the vegas), and S(0) enters once per path through w S(0).
struct AssetConst {
  float w, ws0, mudt, inv_sig, beta;
};

template <int T>
__global__ void __launch_bounds__(kBasketThreads)
    k_basket_asian(const BasketArgs g, const float* __restrict__ z,
                   const float* __restrict__ lt /* (A, A): sigma_a sqrt(dt) L[a][f] */,
                   const float* __restrict__ x0, const float* __restrict__ vols,
                   const float* __restrict__ weights, float rate, double* __restrict__ part) {
  constexpr int K4 = 4 * T;  // factor groups of 4 (16 T factors)
  __shared__ double red[kBasketThreads / 64];
  __shared__ double rsum[(kBasketThreads / 64) * 2 * kMaxBasketAssets];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const int col = lane & 15, row = lane >> 4;
  // B fragments: B[k][j] = sigma_a sqrt(dt) L[a][f], a = 16 t + col, f = 4 kk + row.
  float bf[T][K4];
  AssetConst c[T];
#pragma unroll
  for (int t = 0; t < T; ++t) {
    const int a = 16 * t + col;
    const bool ok = a < g.A;
#pragma unroll
    for (int kk = 0; kk < K4; ++kk) {
      const int f = 4 * kk + row;
      bf[t][kk] = (ok && f <= a && f < g.A) ? lt[a * g.A + f] : 0.0f;
    }
    const float sg = ok ? vols[a] : 1.0f;
    c[t].w = ok ? weights[a] : 0.0f;
    c[t].ws0 = ok ? weights[a] * expf(x0[a]) : 0.0f;
    c[t].mudt = ok ? (rate - 0.5f * sg * sg) * g.dt : 0.0f;
    c[t].inv_sig = 1.0f / sg;
    // dy/dsigma at t_s = (y - (r + sigma^2/2) t_s) / sigma = y inv_sig + beta t_s
    c[t].beta = -(rate + 0.5f * sg * sg) / sg;
  }
  float dacc[T], vacc[T], pacc = 0.0f;
#pragma unroll
  for (int t = 0; t < T; ++t) dacc[t] = vacc[t] = 0.0f;
  const size_t P = static_cast<size_t>(g.P);
  const int ntiles = (g.P + 63) / 64;
  for (int tile = blockIdx.x * nw + wave; tile < ntiles; tile += gridDim.x * nw) {
    const int p4 = tile * 64 + 4 * col;  // this lane's 4 consecutive paths (A operands)
    const bool lane_ok = p4 < g.P;
    float x[4][T][4], D[4][T][4], V[4][T][4];
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int t = 0; t < T; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          x[q][t][r] = 0.0f;
          D[q][t][r] = V[q][t][r] = 0.0f;
        }
    for (int s = 0; s < g.S; ++s) {
      // A operands: factor 4 kk + row of this lane's four paths
      floatx4 za[K4];
#pragma unroll
      for (int kk = 0; kk < K4; ++kk) {
        const int f = 4 * kk + row;
        za[kk] = (lane_ok && f < g.F)
                     ? *reinterpret_cast<const floatx4*>(z + (static_cast<size_t>(s) * g.F + f) * P + p4)
                     : floatx4{0.0f, 0.0f, 0.0f, 0.0f};
      }
      const float ts = (s + 1) * g.dt;
#pragma unroll
      for (int t = 0; t < T; ++t) {
        const float kt = c[t].beta * ts;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          floatx4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
          // factors f <= a only: groups kk < 4 (t + 1) for asset tile t
#pragma unroll
          for (int kk = 0; kk < 4 * (t + 1); ++kk) acc = mfma4(za[kk][q], bf[t][kk], acc);
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const float xv = x[q][t][r] + (c[t].mudt + acc[r]);
            x[q][t][r] = xv;
            const float sv = __expf(xv);
            D[q][t][r] += sv;
            V[q][t][r] += sv * (xv * c[t].inv_sig + kt);
          }
        }
      }
    }
    // Per path: basket average (row reduction over the assets), payoff, and
    // the adjoint weights for this lane's assets.
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int p = tile * 64 + 16 * row + 4 * r + q;
        float b = 0.0f;
#pragma unroll
        for (int t = 0; t < T; ++t) b += c[t].ws0 * D[q][t][r];
        const float avg = row16_sum(b) * g.inv_steps;
        const bool itm = p < g.P && avg > g.strike;
        const float k = itm ? g.disc * g.inv_steps : 0.0f;
        if (col == 0 && itm) pacc += g.disc * (avg - g.strike);
#pragma unroll
        for (int t = 0; t < T; ++t) {
          dacc[t] += k * c[t].w * D[q][t][r];    // d(S0 e^y)/dS0 = e^y
          vacc[t] += k * c[t].ws0 * V[q][t][r];
        }
      }
  }
  // Fold the four rows (same asset, other paths).
#pragma unroll
  for (int t = 0; t < T; ++t) {
    float d = dacc[t], v = vacc[t];
    d += __shfl_xor(d, 16, 64);
    d += __shfl_xor(d, 32, 64);
    v += __shfl_xor(v, 16, 64);
    v += __shfl_xor(v, 32, 64);
    if (row == 0) {
      const int a = 16 * t + col;
      rsum[(wave * 2) * kMaxBasketAssets + a] = d;
      rsum[(wave * 2 + 1) * kMaxBasketAssets + a] = v;
    }
  }
  const int ncols = 1 + 2 * g.A;
  double* out = part + static_cast<size_t>(blockIdx.x) * ncols;
  block_sum_to(static_cast<double>(pacc), red, out);
  // block_sum_to ends with a barrier-free lane-0 store; rsum was written
  // before its barriers.
  for (int a = threadIdx.x; a < g.A; a += blockDim.x) {
    double d = 0.0, v = 0.0;
    for (int w = 0; w < nw; ++w) {
      d += rsum[(w * 2) * kMaxBasketAssets + a];
      v += rsum[(w * 2 + 1) * kMaxBasketAssets + a];
    }
    out[1 + a] = d;
    out[1 + g.A + a] = v;
  }
}

// correlation_kernel_texture_multistep (Correlate_gpu.cu:26-114): per step
// out(s, r, p) = sum_f A(s, r, f) in(s, p, f) for r >= alive(s); lane per path.
__global__ void k_correlate_multistep(const float* __restrict__ in, float* __restrict__ out,
                                      const float* __restrict__ A, const int* __restrict__ alive,
                                      int a_offset_per_step, int factors, int out_dims, int paths,
                                      int steps) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= paths) return;
  const size_t P = static_cast<size_t>(paths);
  for (int s = 0; s < steps; ++s) {
    const float* zin = in + (static_cast<size_t>(s) * paths + p) * factors;
    const float* As = A + static_cast<size_t>(s) * a_offset_per_step;
    float* o = out + static_cast<size_t>(s) * out_dims * P + p;
    for (int r = alive[s]; r < out_dims; ++r) {
      float tot = 0.0f;
      for (int f = 0; f < factors; ++f) tot += As[r * factors + f] * zin[f];
      o[r * P] = tot;
    }
  }
}

// multi_dim_BS_path_generator_kernel (multid_path_gen_BS_gpu.cu:14-64).
__global__ void k_multid_bs_paths(const float* __restrict__ log0, const float* __restrict__ incr,
                                  float* __restrict__ out, int paths, int rates, int steps) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= paths) return;
  const size_t P = static_cast<size_t>(paths);
  for (int r = 0; r < rates; ++r) {
    float x = log0[r];
    for (int s = 0; s < steps; ++s) {
      const size_t o = (static_cast<size_t>(s) * rates + r) * P + p;
      x += incr[o];
      out[o] = expf(x);
    }
  }
}

// basketWeightings_kernel (weighted_average_gpu.cu:14-45).
__global__ void k_basket_weightings(const float* __restrict__ in, float* __restrict__ out,
                                    const float* __restrict__ w, int paths, int stocks, int steps) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= paths) return;
  const size_t P = static_cast<size_t>(paths);
  for (int s = 0; s < steps; ++s) {
    float tot = 0.0f;
    for (int j = 0; j < stocks; ++j) tot += w[j] * in[(static_cast<size_t>(s) * stocks + j) * P + p];
    out[static_cast<size_t>(s) * P + p] = tot;
  }
}

int basket_blocks(int paths) {
  const int tiles = (paths + 63) / 64;
  const int want = (tiles + kBasketThreads / 64 - 1) / (kBasketThreads / 64);
  const int cap = 4 * device_cu_count();
  return want < 1 ? 1 : (want > cap ? cap : want);
}

}  // namespace
}  // namespace koo

using namespace koo;

extern "C" int correlated_paths_multistep_GPU(const float* input_data_device,
                                              float* output_data_device, const float* A_device,
                                              const int* alive_device, int A_offset_per_step,
                                              int factors, int out_dimensions, int number_paths,
                                              int steps, void* stream) {
  if (factors < 1 || out_dimensions < 1 || number_paths < 0 || steps < 1 || A_offset_per_step < 0)
    return set_error(KOO_ERR_INVALID,
                     "correlated_paths_multistep_GPU: bad sizes (factors %d, out %d, paths %d, "
                     "steps %d)",
                     factors, out_dimensions, number_paths, steps);
  if (number_paths == 0) return KOO_OK;
  if (!input_data_device || !output_data_device || !A_device || !alive_device)
    return set_error(KOO_ERR_INVALID, "correlated_paths_multistep_GPU: null pointer");
  hipLaunchKernelGGL(k_correlate_multistep, dim3((number_paths + 255) / 256), dim3(256), 0,
                     as_stream(stream), input_data_device, output_data_device, A_device,
                     alive_device, A_offset_per_step, factors, out_dimensions, number_paths,
                     steps);
  return check_launch("correlated_paths_multistep_GPU");
}

extern "C" int multi_dim_BS_path_generator_kernel(const float* logRates_device,
                                                  const float* correlatedBrownianIncrements_device,
                                                  float* outputPaths_device, int paths, int rates,
                                                  int steps, void* stream) {
  if (paths < 0 || rates < 1 || steps < 1)
    return set_error(KOO_ERR_INVALID, "multi_dim_BS_path_generator_kernel: bad sizes");
  if (paths == 0) return KOO_OK;
  if (!logRates_device || !correlatedBrownianIncrements_device || !outputPaths_device)
    return set_error(KOO_ERR_INVALID, "multi_dim_BS_path_generator_kernel: null pointer");
  hipLaunchKernelGGL(k_multid_bs_paths, dim3((paths + 255) / 256), dim3(256), 0, as_stream(stream),
                     logRates_device, correlatedBrownianIncrements_device, outputPaths_device,
                     paths, rates, steps);
  return check_launch("multi_dim_BS_path_generator_kernel");
}

extern "C" int basketWeightings_gpu(const float* inputPaths, float* outputPaths,
                                    const float* weights, int paths, int numberOfStocks,
                                    int numberSteps, void* stream) {
  if (paths < 0 || numberOfStocks < 1 || numberSteps < 1)
    return set_error(KOO_ERR_INVALID, "basketWeightings_gpu: bad sizes");
  if (paths == 0) return KOO_OK;
  if (!inputPaths || !outputPaths || !weights)
    return set_error(KOO_ERR_INVALID, "basketWeightings_gpu: null pointer");
  hipLaunchKernelGGL(k_basket_weightings, dim3((paths + 255) / 256), dim3(256), 0,
                     as_stream(stream), inputPaths, outputPaths, weights, paths, numberOfStocks,
                     numberSteps);
  return check_launch("basketWeightings_gpu");
}

extern "C" int basket_asian_workspace(int paths, int assets, size_t* partial_bytes) {
  if (paths < 1 || assets < 1 || assets > kMaxBasketAssets)
    return set_error(KOO_ERR_INVALID, "basket_asian_workspace: need paths>=1, 1<=assets<=%d",
                     kMaxBasketAssets);
  if (partial_bytes)
    *partial_bytes = sizeof(double) * static_cast<size_t>(basket_blocks(paths)) * (1 + 2 * assets);
  return KOO_OK;
}

extern "C" int basket_asian_gpu(int paths, int assets, int steps, const float* variates,
                                const float* log_spots, const float* vols,
                                const float* scaled_chol, const float* weights, float rate,
                                float dt, float strike, double* partials, double* out,
                                void* stream) {
  if (paths < 4 || paths % 4 != 0 || assets < 1 || assets > kMaxBasketAssets || steps < 1 ||
      !(dt > 0.0f))
    return set_error(KOO_ERR_INVALID,
                     "basket_asian_gpu: need paths a positive multiple of 4, 1<=assets<=%d, "
                     "steps>=1, dt>0 (got %d, %d, %d)",
                     kMaxBasketAssets, paths, assets, steps);
  if (!variates || !log_spots || !vols || !scaled_chol || !weights || !partials || !out)
    return set_error(KOO_ERR_INVALID, "basket_asian_gpu: null pointer");
  if (reinterpret_cast<uintptr_t>(variates) % 16 != 0)
    return set_error(KOO_ERR_INVALID, "basket_asian_gpu: variates must be 16-byte aligned");
  BasketArgs g;
  g.P = paths;
  g.A = assets;
  g.S = steps;
  g.F = assets;
  g.dt = dt;
  g.strike = strike;
  g.disc = expf(-rate * dt * steps);
  g.inv_steps = 1.0f / static_cast<float>(steps);
  const int nb = basket_blocks(paths);
  hipStream_t st = as_stream(stream);
  switch ((assets + 15) / 16) {
    case 1:
      hipLaunchKernelGGL(k_basket_asian<1>, dim3(nb), dim3(kBasketThreads), 0, st, g, variates,
                         scaled_chol, log_spots, vols, weights, rate, partials);
      break;
    case 2:
      hipLaunchKernelGGL(k_basket_asian<2>, dim3(nb), dim3(kBasketThreads), 0, st, g, variates,
                         scaled_chol, log_spots, vols, weights, rate, partials);
      break;
    case 3:
      hipLaunchKernelGGL(k_basket_asian<3>, dim3(nb), dim3(kBasketThreads), 0, st, g, variates,
                         scaled_chol, log_spots, vols, weights, rate, partials);
      break;
    default:
      hipLaunchKernelGGL(k_basket_asian<4>, dim3(nb), dim3(kBasketThreads), 0, st, g, variates,
                         scaled_chol, log_spots, vols, weights, rate, partials);
      break;
  }
  hipLaunchKernelGGL(k_finalize_cols, dim3(1 + 2 * assets), dim3(256), 0, st, partials, nb,
                     1 + 2 * assets, 1.0 / static_cast<double>(paths), out);
  return check_launch("basket_asian_gpu");
}
