// lmm.hip -- stage (b)/(c) evolution kernels behind the reference's
// LMM_evolver_* / *_computation_gpu entry points, for gfx950.
//
// One lane per path.  Model data that is the same for every path (pseudo-
// roots, taus, displacements, drifts, alive indices) is read with uniform
// addresses, so it goes through the scalar unit; per-path data is path-minor
// and every wavefront access is a contiguous row.  The predictor-corrector
// evolver keeps the running drift sums e_f, e^_f in registers (factors is a
// template parameter) and the per-path rate state in LDS, so the only HBM
// traffic is the variates read and the rate / log-rate / discount cubes
// written.
#include "koo_common.h"
#include "lmm_step.h"

namespace koo {

namespace {

using lmm::fast_exp;
using lmm::fast_rcp;

constexpr int kMaxSteps = 512;
constexpr int kMaxRates = 128;
constexpr int kMaxFactors = 8;

// A host index vector (the reference's std::vector<int>) passed by value.
struct IndexVec {
  int v[kMaxSteps];
};

inline IndexVec to_index_vec(const int* host, int n) {
  IndexVec iv{};
  for (int i = 0; i < n; ++i) iv.v[i] = host[i];
  return iv;
}

inline int lanes_for_state(int floats_per_path) {
  // Largest multiple of 64 lanes (<= 256) whose LDS state fits 64 KiB.
  int w = 65536 / (static_cast<int>(sizeof(float)) * floats_per_path * 64);
  if (w > 4) w = 4;
  if (w < 1) w = 1;
  return 64 * w;
}

// ----------------------------------------------------------------------
// correlated_drift_paths_GPU: y(s,r,p) = sum_f A(s,r,f) z(s,p,f) + mu(s,r)
// for r >= alive(s) (correlate_drift_gpu.cu:142-220).
// ----------------------------------------------------------------------
template <int F>
__global__ void __launch_bounds__(256)
    k_correlate_drift(const float* __restrict__ in, float* __restrict__ out,
                      const float* __restrict__ A, const int* __restrict__ alive,
                      const float* __restrict__ drift, int a_off, int out_dims, int paths,
                      int steps) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= paths) return;
  for (int s = 0; s < steps; ++s) {
    float z[F];
#pragma unroll
    for (int f = 0; f < F; ++f) z[f] = in[(static_cast<size_t>(s) * paths + p) * F + f];
    const float* As = A + static_cast<size_t>(s) * a_off;
    for (int r = alive[s]; r < out_dims; ++r) {
      float t = 0.0f;
#pragma unroll
      for (int f = 0; f < F; ++f) t += As[r * F + f] * z[f];
      out[(static_cast<size_t>(s) * out_dims + r) * paths + p] = t + drift[s * out_dims + r];
    }
  }
}

// ----------------------------------------------------------------------
// LMM_evolver_pc_all_gpu: predictor-corrector log-Euler over all steps,
// discount ratios fused (LMM_evolver_all.cu:506-747).  Step 0 takes the host
// initial drifts; later steps recompute the drifts from the previous step's
// rates: the log-rates stay in LDS ([rate][lane]) and f + d = exp(log-rate)
// is recomputed (one v_exp_f32 instead of a second LDS curve, so twice the
// lanes fit per CU).  exp and 1/x are the hardware forms (v_exp_f32,
// v_rcp_f32, about 1 ulp), as in the fused pricing kernels.
// ----------------------------------------------------------------------
template <int F>
__global__ void __launch_bounds__(256)
    k_lmm_pc_all(const float* __restrict__ f0, const float* __restrict__ logf0,
                 const float* __restrict__ taus, const float* __restrict__ variates,
                 const float* __restrict__ A, const float* __restrict__ mu0,
                 const float* __restrict__ fixed, const float* __restrict__ disp,
                 const int* __restrict__ alive, int paths, int steps, int rates, int path_mult,
                 int fact_mult, float* __restrict__ rates_out, float* __restrict__ logs_out,
                 float* __restrict__ disc_out, int do_disc) {
  extern __shared__ float lds[];
  const int bd = blockDim.x;
  float* L = lds + threadIdx.x;  // L[r * bd]: log(f + d) of the curve
  const int p = blockIdx.x * bd + threadIdx.x;
  if (p >= paths) return;
  const size_t P = static_cast<size_t>(paths);
  (void)f0;  // step 0 takes the host drifts; later steps rebuild f + d = exp(L)
  for (int r = alive[0]; r < rates; ++r) L[r * bd] = logf0[r];
  for (int s = 0; s < steps; ++s) {
    const int a = alive[s];
    float z[F], e[F], ep[F];
    const float* vs = variates + static_cast<size_t>(s) * F * P + static_cast<size_t>(p) * path_mult;
#pragma unroll
    for (int f = 0; f < F; ++f) {
      z[f] = vs[static_cast<size_t>(f) * fact_mult];
      e[f] = 0.0f;
      ep[f] = 0.0f;
    }
    const float* As = A + static_cast<size_t>(s) * rates * F;
    float df = 1.0f;
    float* dcol = disc_out + static_cast<size_t>(s) * (rates + 1) * P + p;
    if (do_disc) dcol[static_cast<size_t>(a) * P] = 1.0f;
    for (int r = a; r < rates; ++r) {
      const float tau = taus[r], d = disp[r];
      float ar[F];
#pragma unroll
      for (int f = 0; f < F; ++f) ar[f] = As[r * F + f];
      float mu;
      if (s == 0) {
        mu = mu0[r];
      } else {
        const float x = fast_exp(L[r * bd]);  // f + d at the start of the step
        const float g = tau * x * fast_rcp(1.0f + tau * (x - d));
        mu = 0.0f;
#pragma unroll
        for (int f = 0; f < F; ++f) {
          e[f] += g * ar[f];
          mu += ar[f] * e[f];
        }
      }
      float dW = 0.0f;
#pragma unroll
      for (int f = 0; f < F; ++f) dW += ar[f] * z[f];
      float lf = L[r * bd] + mu + dW + fixed[s * rates + r];
      const float xp = fast_exp(lf);
      const float gp = tau * xp * fast_rcp(1.0f + tau * (xp - d));
      float mup = 0.0f;
#pragma unroll
      for (int f = 0; f < F; ++f) {
        ep[f] += gp * ar[f];
        mup += ar[f] * ep[f];
      }
      lf += 0.5f * (mup - mu);
      const float rate = fast_exp(lf) - d;
      L[r * bd] = lf;
      const size_t o = (static_cast<size_t>(s) * rates + r) * P + p;
      logs_out[o] = lf;
      rates_out[o] = rate;
      if (do_disc) {
        df *= fast_rcp(1.0f + rate * tau);
        dcol[static_cast<size_t>(r + 1) * P] = df;
      }
    }
  }
}

// ----------------------------------------------------------------------
// LMM_evolver_pc_gpu / LMM_evolver_euler_gpu on correlated increments
// incr (S,N,P) = A z + fixed drift.  The reference launches one kernel per
// step and reads the previous step's rate and log-rate cubes back
// (LMM_evolver_gpu.cu:366-833, 835-1180); here one launch walks every step
// with the curve in LDS ([rate][lane]) and only writes the cubes.  Step 0
// takes the host initial drifts for every rate; step s >= 1 rebuilds the
// drift over r >= alive[s] from the previous step's rates,
// g = (f tau + d tau) / (1 + f tau).  PC adds the predictor-corrector half
// step; Euler caps the log of the displaced rate at log 10
// (LMM_evolver_gpu.cu:20).
// ----------------------------------------------------------------------
constexpr float kLogCeiling = 2.302585092994045684f;  // log 10

template <int F, bool PC>
__global__ void __launch_bounds__(256)
    k_lmm_steps(const float* __restrict__ f0, const float* __restrict__ logf0,
                const float* __restrict__ taus, const float* __restrict__ incr,
                const float* __restrict__ A, const float* __restrict__ mu0,
                const float* __restrict__ disp, const IndexVec alive, int paths, int steps,
                int rates, float* __restrict__ rates_out, float* __restrict__ logs_out) {
  extern __shared__ float lds[];
  const int bd = blockDim.x;
  float* L = lds + threadIdx.x;               // L[r * bd]
  float* R = lds + rates * bd + threadIdx.x;  // R[r * bd]
  const int p = blockIdx.x * bd + threadIdx.x;
  if (p >= paths) return;
  const size_t P = static_cast<size_t>(paths);
  for (int r = 0; r < rates; ++r) {
    L[r * bd] = logf0[r];
    R[r * bd] = f0[r];
  }
  for (int s = 0; s < steps; ++s) {
    const int a = s == 0 ? 0 : alive.v[s];
    float e[F], ep[F];
#pragma unroll
    for (int f = 0; f < F; ++f) e[f] = ep[f] = 0.0f;
    const float* As = A + static_cast<size_t>(s) * rates * F;
    const float* in = incr + static_cast<size_t>(s) * rates * P + p;
    for (int r = a; r < rates; ++r) {
      const float tau = taus[r], d = disp[r];
      float ar[F];
#pragma unroll
      for (int f = 0; f < F; ++f) ar[f] = As[r * F + f];
      float mu;
      if (s == 0) {
        mu = mu0[r];
      } else {
        const float fr = R[r * bd];
        const float frt = fr * tau;
        const float g = (frt + d * tau) / (1.0f + frt);
        mu = 0.0f;
#pragma unroll
        for (int f = 0; f < F; ++f) {
          e[f] += g * ar[f];
          mu += e[f] * ar[f];
        }
      }
      float lf = L[r * bd] + (mu + in[static_cast<size_t>(r) * P]);
      if (PC) {
        const float xp = expf(lf);
        const float gp = xp * tau / (1.0f + (xp - d) * tau);
        float mup = 0.0f;
#pragma unroll
        for (int f = 0; f < F; ++f) {
          ep[f] += gp * ar[f];
          mup += ar[f] * ep[f];
        }
        lf += 0.5f * (mup - mu);
      } else if (lf > kLogCeiling) {
        lf = kLogCeiling;
      }
      const float rate = expf(lf) - d;
      L[r * bd] = lf;
      R[r * bd] = rate;
      const size_t o = (static_cast<size_t>(s) * rates + r) * P + p;
      logs_out[o] = lf;
      rates_out[o] = rate;
    }
  }
}

// P(s,alive) = 1, P(s,r+1) = P(s,r)/(1+tau_r f(s,r))  (LMM_evolver_gpu.cu:1188-1293).
__global__ void k_discount_ratios(const float* __restrict__ rates, const float* __restrict__ taus,
                                  const IndexVec alive, int paths, int steps, int n_rates,
                                  float* __restrict__ disc) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= paths) return;
  const size_t P = static_cast<size_t>(paths);
  for (int s = 0; s < steps; ++s) {
    const int a = alive.v[s];
    float* dcol = disc + static_cast<size_t>(s) * (n_rates + 1) * P + p;
    const float* rcol = rates + static_cast<size_t>(s) * n_rates * P + p;
    float df = 1.0f;
    dcol[static_cast<size_t>(a) * P] = df;
    for (int r = a; r < n_rates; ++r) {
      df /= 1.0f + taus[r] * rcol[static_cast<size_t>(r) * P];
      dcol[static_cast<size_t>(r + 1) * P] = df;
    }
  }
}

// Ann(s,r) = sum_{j>=r} tau_j P(s,j+1)  (LMM_evolver_gpu.cu:1415-1461).
__global__ void k_annuities(const float* __restrict__ disc, const float* __restrict__ taus,
                            const IndexVec alive, int paths, int steps, int n_rates,
                            float* __restrict__ ann) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= paths) return;
  const size_t P = static_cast<size_t>(paths);
  for (int s = 0; s < steps; ++s) {
    const float* dcol = disc + static_cast<size_t>(s) * (n_rates + 1) * P + p;
    float* acol = ann + static_cast<size_t>(s) * n_rates * P + p;
    float a = 0.0f;
    for (int r = n_rates - 1; r >= alive.v[s]; --r) {
      a += taus[r] * dcol[static_cast<size_t>(r + 1) * P];
      acol[static_cast<size_t>(r) * P] = a;
    }
  }
}

// SR(s,r) = (P(s,r) - P(s,N)) / Ann(s,r)  (LMM_evolver_gpu.cu:1548-1595).
__global__ void k_cot_swap_rates(const float* __restrict__ disc, const float* __restrict__ ann,
                                 const IndexVec alive, int paths, int steps, int n_rates,
                                 float* __restrict__ sr) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= paths) return;
  const size_t P = static_cast<size_t>(paths);
  for (int s = 0; s < steps; ++s) {
    const float* dcol = disc + static_cast<size_t>(s) * (n_rates + 1) * P + p;
    const float* acol = ann + static_cast<size_t>(s) * n_rates * P + p;
    float* scol = sr + static_cast<size_t>(s) * n_rates * P + p;
    const float df2 = dcol[static_cast<size_t>(n_rates) * P];
    for (int r = alive.v[s]; r < n_rates; ++r)
      scol[static_cast<size_t>(r) * P] =
          (dcol[static_cast<size_t>(r) * P] - df2) / acol[static_cast<size_t>(r) * P];
  }
}

// Swap rate on one step's curve  (LMM_evolver_gpu.cu:1655-1709).
__global__ void k_swap_rate(const float* __restrict__ disc, const float* __restrict__ taus,
                            int start, int end, int paths, int step_in, int step_out,
                            int n_rates, float* __restrict__ out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= paths) return;
  const size_t P = static_cast<size_t>(paths);
  const float* dcol = disc + static_cast<size_t>(step_in) * (n_rates + 1) * P + p;
  const float num = dcol[static_cast<size_t>(start) * P] - dcol[static_cast<size_t>(end) * P];
  float den = 0.0f;
  for (int r = start; r < end; ++r) den += taus[r] * dcol[static_cast<size_t>(r + 1) * P];
  out[static_cast<size_t>(step_out) * P + p] = num / den;
}

// B(0) = 1, B(i) = B(i-1) P(i-1,i-1)/P(i-1,i)  (LMM_evolver_gpu.cu:1785-1838).
// Four paths per lane (16-byte loads and stores) when the path counts and
// the offset keep every row 16-byte aligned; the ratios' loads do not depend
// on the running product, so they are all issued ahead of it.
__global__ void __launch_bounds__(256)
    k_numeraires_vec4(const f32x4* __restrict__ disc, f32x4* __restrict__ num, int paths4,
                      int out_paths4, int out_offset4, int n_rates) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= paths4) return;
  const size_t P = static_cast<size_t>(paths4), Q = static_cast<size_t>(out_paths4);
  f32x4 b = {1.0f, 1.0f, 1.0f, 1.0f};
  store_stream(b, num + out_offset4 + p);
  for (int i = 1; i < n_rates; ++i) {
    const f32x4* dcol = disc + static_cast<size_t>(i - 1) * (n_rates + 1) * P + p;
    const f32x4 a = load_stream(dcol + static_cast<size_t>(i - 1) * P);
    const f32x4 c = load_stream(dcol + static_cast<size_t>(i) * P);
    b *= a / c;
    store_stream(b, num + static_cast<size_t>(i) * Q + out_offset4 + p);
  }
}

__global__ void k_numeraires(const float* __restrict__ disc, float* __restrict__ num, int paths,
                             int out_paths, int out_offset, int n_rates) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= paths) return;
  const size_t P = static_cast<size_t>(paths), Q = static_cast<size_t>(out_paths);
  float b = 1.0f;
  num[out_offset + p] = b;
  for (int i = 1; i < n_rates; ++i) {
    const float* dcol = disc + static_cast<size_t>(i - 1) * (n_rates + 1) * P + p;
    b *= dcol[static_cast<size_t>(i - 1) * P] / dcol[static_cast<size_t>(i) * P];
    num[static_cast<size_t>(i) * Q + out_offset + p] = b;
  }
}

// out(i,p) = f(step[i], index[i], p)  (LMM_evolver_gpu.cu:1910-2038): a row
// copy per output row, so the 16-byte form has one row per blockIdx.y and
// four paths per lane, with streaming (nontemporal) loads and stores.
__global__ void __launch_bounds__(256)
    k_forward_extract_vec4(const f32x4* __restrict__ fwd, const IndexVec index,
                           const IndexVec step, int paths4, int n_rates, f32x4* __restrict__ out) {
  const int i = blockIdx.y;
  const size_t P4 = static_cast<size_t>(paths4);
  const f32x4* src = fwd + (static_cast<size_t>(step.v[i]) * n_rates + index.v[i]) * P4;
  f32x4* dst = out + static_cast<size_t>(i) * P4;
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < paths4; p += gridDim.x * blockDim.x)
    store_stream(load_stream(src + p), dst + p);
}

__global__ void k_forward_extract(const float* __restrict__ fwd, const IndexVec index,
                                  const IndexVec step, int n_out, int paths, int n_rates,
                                  float* __restrict__ out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= paths) return;
  const size_t P = static_cast<size_t>(paths);
  for (int i = 0; i < n_out; ++i)
    out[static_cast<size_t>(i) * P + p] =
        fwd[(static_cast<size_t>(step.v[i]) * n_rates + index.v[i]) * P + p];
}

// Numeraires at exercise dates  (LMM_evolver_gpu.cu:2095-2141).
__global__ void k_numeraire_extract(const float* __restrict__ all, float* __restrict__ some,
                                    int paths_in, int paths_out, int offset, int n_ex,
                                    const int* __restrict__ ex) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= paths_in) return;
  for (int i = 0; i < n_ex; ++i)
    some[offset + p + static_cast<size_t>(paths_out) * i] =
        all[p + static_cast<size_t>(paths_in) * ex[i]];
}

inline dim3 path_grid(int paths, int threads) { return dim3((paths + threads - 1) / threads); }

}  // namespace
}  // namespace koo

using namespace koo;

extern "C" int correlated_drift_paths_GPU(float* in, float* out, float* A, int* alive,
                                          float* drift, int A_offset_per_step, int factors,
                                          int out_dimensions, int number_paths, int steps,
                                          void* stream) {
  if (factors < 1 || factors > kMaxFactors || number_paths < 0 || steps < 0 || out_dimensions < 1)
    return set_error(KOO_ERR_INVALID, "correlated_drift_paths_GPU: bad sizes (F=%d)", factors);
  if (number_paths == 0 || steps == 0) return KOO_OK;
  if (!in || !out || !A || !alive || !drift)
    return set_error(KOO_ERR_INVALID, "correlated_drift_paths_GPU: null pointer");
  dim3 g = path_grid(number_paths, 256);
  hipStream_t st = as_stream(stream);
  switch (factors) {
#define KOO_CASE(F)                                                                          \
  case F:                                                                                    \
    hipLaunchKernelGGL(k_correlate_drift<F>, g, dim3(256), 0, st, in, out, A, alive, drift, \
                       A_offset_per_step, out_dimensions, number_paths, steps);             \
    break;
    KOO_CASE(1) KOO_CASE(2) KOO_CASE(3) KOO_CASE(4) KOO_CASE(5) KOO_CASE(6) KOO_CASE(7)
    KOO_CASE(8)
#undef KOO_CASE
  }
  return check_launch("correlated_drift_paths_GPU");
}

static int check_alive(const char* who, const int* alive_host, int steps, int n_rates) {
  if (steps < 0 || steps > kMaxSteps)
    return set_error(KOO_ERR_INVALID, "%s: steps %d outside [0,%d]", who, steps, kMaxSteps);
  if (!alive_host && steps > 0) return set_error(KOO_ERR_INVALID, "%s: null alive indices", who);
  for (int s = 0; s < steps; ++s)
    if (alive_host[s] < 0 || alive_host[s] > n_rates)
      return set_error(KOO_ERR_INVALID, "%s: alive[%d] = %d outside [0,%d]", who, s, alive_host[s],
                       n_rates);
  return KOO_OK;
}

extern "C" int LMM_evolver_pc_all_gpu(float* initial_rates, float* initial_log_rates,
                                      float* taus, float* variates, float* pseudo_roots,
                                      float* initial_drifts, float* fixed_drifts,
                                      float* displacements, int* alive, int paths, int factors,
                                      int steps, int rates, float* /*e_buffer*/,
                                      float* /*e_buffer_pred*/, float* evolved_rates,
                                      float* evolved_log_rates, float* discounts,
                                      int /*sharedMemForLogs*/, int transposeData, void* stream,
                                      int /*numberThreads*/, int doDiscounts) {
  if (factors < 1 || factors > kMaxFactors || rates < 1 || rates > kMaxRates || paths < 0 ||
      steps < 0)
    return set_error(KOO_ERR_INVALID,
                     "LMM_evolver_pc_all_gpu: need 1<=factors<=%d, 1<=rates<=%d (got %d, %d)",
                     kMaxFactors, kMaxRates, factors, rates);
  if (paths == 0 || steps == 0) return KOO_OK;
  if (!initial_rates || !initial_log_rates || !taus || !variates || !pseudo_roots ||
      !initial_drifts || !fixed_drifts || !displacements || !alive || !evolved_rates ||
      !evolved_log_rates || (doDiscounts && !discounts))
    return set_error(KOO_ERR_INVALID, "LMM_evolver_pc_all_gpu: null pointer");
  const int path_mult = transposeData ? 1 : factors;
  const int fact_mult = transposeData ? paths : 1;
  const int threads = lanes_for_state(rates);
  const size_t lds = sizeof(float) * rates * threads;
  dim3 g = path_grid(paths, threads);
  hipStream_t st = as_stream(stream);
  switch (factors) {
#define KOO_CASE(F)                                                                             \
  case F:                                                                                       \
    hipLaunchKernelGGL(k_lmm_pc_all<F>, g, dim3(threads), lds, st, initial_rates,               \
                       initial_log_rates, taus, variates, pseudo_roots, initial_drifts,         \
                       fixed_drifts, displacements, alive, paths, steps, rates, path_mult,      \
                       fact_mult, evolved_rates, evolved_log_rates, discounts, doDiscounts);    \
    break;
    KOO_CASE(1) KOO_CASE(2) KOO_CASE(3) KOO_CASE(4) KOO_CASE(5) KOO_CASE(6) KOO_CASE(7)
    KOO_CASE(8)
#undef KOO_CASE
  }
  return check_launch("LMM_evolver_pc_all_gpu");
}

static int lmm_steps_launch(const char* who, bool pc, float* initial_rates,
                            float* initial_log_rates, float* taus, float* increments,
                            float* pseudo_roots, float* initial_drifts, float* displacements,
                            const int* alive_host, int paths, int factors, int steps, int rates,
                            float* evolved_rates, float* evolved_log_rates, void* stream) {
  if (factors < 1 || factors > kMaxFactors || rates < 1 || rates > kMaxRates || paths < 0)
    return set_error(KOO_ERR_INVALID, "%s: need 1<=factors<=%d, 1<=rates<=%d (got %d, %d)", who,
                     kMaxFactors, kMaxRates, factors, rates);
  if (int e = check_alive(who, alive_host, steps, rates)) return e;
  if (paths == 0 || steps == 0) return KOO_OK;
  if (!initial_rates || !initial_log_rates || !taus || !increments || !pseudo_roots ||
      !initial_drifts || !displacements || !evolved_rates || !evolved_log_rates)
    return set_error(KOO_ERR_INVALID, "%s: null pointer", who);
  const int threads = lanes_for_state(2 * rates);
  const size_t lds = sizeof(float) * 2 * rates * threads;
  const IndexVec alive = to_index_vec(alive_host, steps);
  dim3 g = path_grid(paths, threads);
  hipStream_t st = as_stream(stream);
  switch (factors) {
#define KOO_CASE(F)                                                                            \
  case F:                                                                                      \
    if (pc)                                                                                    \
      hipLaunchKernelGGL((k_lmm_steps<F, true>), g, dim3(threads), lds, st, initial_rates,     \
                         initial_log_rates, taus, increments, pseudo_roots, initial_drifts,    \
                         displacements, alive, paths, steps, rates, evolved_rates,             \
                         evolved_log_rates);                                                   \
    else                                                                                       \
      hipLaunchKernelGGL((k_lmm_steps<F, false>), g, dim3(threads), lds, st, initial_rates,    \
                         initial_log_rates, taus, increments, pseudo_roots, initial_drifts,    \
                         displacements, alive, paths, steps, rates, evolved_rates,             \
                         evolved_log_rates);                                                   \
    break;
    KOO_CASE(1) KOO_CASE(2) KOO_CASE(3) KOO_CASE(4) KOO_CASE(5) KOO_CASE(6) KOO_CASE(7)
    KOO_CASE(8)
#undef KOO_CASE
  }
  return check_launch(who);
}

extern "C" int LMM_evolver_pc_gpu(float* initial_rates, float* initial_log_rates, float* taus,
                                  float* correlated_increments, float* pseudo_roots,
                                  float* initial_drifts, float* displacements,
                                  const int* alive_indices_host, int paths, int factors, int steps,
                                  int rates, float* /*e_buffer*/, float* /*e_buffer_pred*/,
                                  float* evolved_rates, float* evolved_log_rates,
                                  int /*useShared*/, void* stream) {
  return lmm_steps_launch("LMM_evolver_pc_gpu", true, initial_rates, initial_log_rates, taus,
                          correlated_increments, pseudo_roots, initial_drifts, displacements,
                          alive_indices_host, paths, factors, steps, rates, evolved_rates,
                          evolved_log_rates, stream);
}

extern "C" int LMM_evolver_euler_gpu(float* initial_rates, float* initial_log_rates, float* taus,
                                     float* correlated_increments, float* pseudo_roots,
                                     float* initial_drifts, float* displacements,
                                     const int* alive_indices_host, int paths, int factors,
                                     int steps, int rates, float* /*e_buffer*/,
                                     float* evolved_rates, float* evolved_log_rates,
                                     void* stream) {
  return lmm_steps_launch("LMM_evolver_euler_gpu", false, initial_rates, initial_log_rates, taus,
                          correlated_increments, pseudo_roots, initial_drifts, displacements,
                          alive_indices_host, paths, factors, steps, rates, evolved_rates,
                          evolved_log_rates, stream);
}


extern "C" int discount_ratios_computation_gpu(float* evolved_rates, float* taus,
                                               const int* alive_host, int* /*alive_device*/,
                                               int paths, int steps, int rates, float* discounts,
                                               int /*allStepsAtOnce*/, void* stream) {
  if (int e = check_alive("discount_ratios_computation_gpu", alive_host, steps, rates)) return e;
  if (paths <= 0 || steps == 0) return paths < 0 ? set_error(KOO_ERR_INVALID, "paths < 0") : KOO_OK;
  if (!evolved_rates || !taus || !discounts)
    return set_error(KOO_ERR_INVALID, "discount_ratios_computation_gpu: null pointer");
  hipLaunchKernelGGL(k_discount_ratios, path_grid(paths, 256), dim3(256), 0, as_stream(stream),
                     evolved_rates, taus, to_index_vec(alive_host, steps), paths, steps, rates,
                     discounts);
  return check_launch("discount_ratios_computation_gpu");
}

extern "C" int coterminal_annuity_ratios_computation_gpu(float* discounts, float* taus,
                                                         const int* alive_host, int paths,
                                                         int steps, int rates, float* annuities,
                                                         void* stream) {
  if (int e = check_alive("coterminal_annuity_ratios_computation_gpu", alive_host, steps, rates))
    return e;
  if (paths <= 0 || steps == 0) return paths < 0 ? set_error(KOO_ERR_INVALID, "paths < 0") : KOO_OK;
  if (!discounts || !taus || !annuities)
    return set_error(KOO_ERR_INVALID, "coterminal_annuity_ratios_computation_gpu: null pointer");
  hipLaunchKernelGGL(k_annuities, path_grid(paths, 256), dim3(256), 0, as_stream(stream),
                     discounts, taus, to_index_vec(alive_host, steps), paths, steps, rates,
                     annuities);
  return check_launch("coterminal_annuity_ratios_computation_gpu");
}

extern "C" int coterminal_swap_rate_computation_gpu(float* discounts, float* annuities,
                                                    const int* alive_host, int paths, int steps,
                                                    int rates, float* swap_rates, void* stream) {
  if (int e = check_alive("coterminal_swap_rate_computation_gpu", alive_host, steps, rates))
    return e;
  if (paths <= 0 || steps == 0) return paths < 0 ? set_error(KOO_ERR_INVALID, "paths < 0") : KOO_OK;
  if (!discounts || !annuities || !swap_rates)
    return set_error(KOO_ERR_INVALID, "coterminal_swap_rate_computation_gpu: null pointer");
  hipLaunchKernelGGL(k_cot_swap_rates, path_grid(paths, 256), dim3(256), 0, as_stream(stream),
                     discounts, annuities, to_index_vec(alive_host, steps), paths, steps, rates,
                     swap_rates);
  return check_launch("coterminal_swap_rate_computation_gpu");
}

extern "C" int swap_rate_computation_gpu(float* discounts, float* taus, int startIndex,
                                         int endIndex, int paths, int step_in, int step_out,
                                         int rates, float* swap_rate, void* stream) {
  if (startIndex < 0 || endIndex > rates || startIndex >= endIndex || paths < 0 || step_in < 0 ||
      step_out < 0)
    return set_error(KOO_ERR_INVALID, "swap_rate_computation_gpu: bad indices");
  if (paths == 0) return KOO_OK;
  if (!discounts || !taus || !swap_rate)
    return set_error(KOO_ERR_INVALID, "swap_rate_computation_gpu: null pointer");
  hipLaunchKernelGGL(k_swap_rate, path_grid(paths, 256), dim3(256), 0, as_stream(stream),
                     discounts, taus, startIndex, endIndex, paths, step_in, step_out, rates,
                     swap_rate);
  return check_launch("swap_rate_computation_gpu");
}

extern "C" int spot_measure_numeraires_computation_offset_gpu(float* discounts, float* numeraires,
                                                              int paths, int pathsForOutput,
                                                              int pathsOffsetForOutput,
                                                              int n_rates, void* stream) {
  if (paths < 0 || n_rates < 1 || pathsOffsetForOutput < 0 ||
      pathsOffsetForOutput + paths > pathsForOutput)
    return set_error(KOO_ERR_INVALID, "spot_measure_numeraires_computation: bad sizes");
  if (paths == 0) return KOO_OK;
  if (!discounts || !numeraires)
    return set_error(KOO_ERR_INVALID, "spot_measure_numeraires_computation: null pointer");
  if (paths % 4 == 0 && pathsForOutput % 4 == 0 && pathsOffsetForOutput % 4 == 0 &&
      (reinterpret_cast<uintptr_t>(discounts) & 15u) == 0 &&
      (reinterpret_cast<uintptr_t>(numeraires) & 15u) == 0) {
    hipLaunchKernelGGL(k_numeraires_vec4, path_grid(paths / 4, 256), dim3(256), 0,
                       as_stream(stream), reinterpret_cast<const f32x4*>(discounts),
                       reinterpret_cast<f32x4*>(numeraires), paths / 4, pathsForOutput / 4,
                       pathsOffsetForOutput / 4, n_rates);
  } else {
    hipLaunchKernelGGL(k_numeraires, path_grid(paths, 256), dim3(256), 0, as_stream(stream),
                       discounts, numeraires, paths, pathsForOutput, pathsOffsetForOutput,
                       n_rates);
  }
  return check_launch("spot_measure_numeraires_computation");
}

extern "C" int spot_measure_numeraires_computation_gpu(float* discounts, float* numeraires,
                                                       int paths, int n_rates, void* stream) {
  return spot_measure_numeraires_computation_offset_gpu(discounts, numeraires, paths, paths, 0,
                                                        n_rates, stream);
}

extern "C" int forward_rate_extraction_gpu(float* all_forwards, const int* forwardIndices_host,
                                           int paths, int steps, int rates, float* out,
                                           void* stream) {
  if (steps < 0 || steps > kMaxSteps || paths < 0 || !forwardIndices_host)
    return set_error(KOO_ERR_INVALID, "forward_rate_extraction_gpu: bad arguments");
  IndexVec idx{}, st{};
  for (int s = 0; s < steps; ++s) {
    if (forwardIndices_host[s] < 0 || forwardIndices_host[s] >= rates)
      return set_error(KOO_ERR_INVALID, "forward_rate_extraction_gpu: index %d out of range",
                       forwardIndices_host[s]);
    idx.v[s] = forwardIndices_host[s];
    st.v[s] = s;
  }
  if (paths == 0 || steps == 0) return KOO_OK;
  if (!all_forwards || !out) return set_error(KOO_ERR_INVALID, "forward_rate_extraction_gpu: null");
  if (paths % 4 == 0 && (reinterpret_cast<uintptr_t>(all_forwards) & 15u) == 0 &&
      (reinterpret_cast<uintptr_t>(out) & 15u) == 0) {
    const int p4 = paths / 4;
    const unsigned gx = grid_for(p4, 256, 512);  // at most 512 blocks per row
    hipLaunchKernelGGL(k_forward_extract_vec4, dim3(gx, steps), dim3(256), 0, as_stream(stream),
                       reinterpret_cast<const f32x4*>(all_forwards), idx, st, p4, rates,
                       reinterpret_cast<f32x4*>(out));
  } else {
    hipLaunchKernelGGL(k_forward_extract, path_grid(paths, 256), dim3(256), 0, as_stream(stream),
                       all_forwards, idx, st, steps, paths, rates, out);
  }
  return check_launch("forward_rate_extraction_gpu");
}

extern "C" int forward_rate_extraction_selecting_step_gpu(float* all_forwards,
                                                          const int* forwardIndices_host,
                                                          const int* stepIndices_host,
                                                          int n_indices, int paths, int rates,
                                                          float* out, void* stream) {
  if (n_indices < 0 || n_indices > kMaxSteps || paths < 0 || !forwardIndices_host ||
      !stepIndices_host)
    return set_error(KOO_ERR_INVALID, "forward_rate_extraction_selecting_step_gpu: bad args");
  IndexVec idx{}, st{};
  for (int s = 0; s < n_indices; ++s) {
    if (forwardIndices_host[s] < 0 || forwardIndices_host[s] >= rates || stepIndices_host[s] < 0)
      return set_error(KOO_ERR_INVALID, "forward_rate_extraction_selecting_step_gpu: index");
    idx.v[s] = forwardIndices_host[s];
    st.v[s] = stepIndices_host[s];
  }
  if (paths == 0 || n_indices == 0) return KOO_OK;
  if (!all_forwards || !out)
    return set_error(KOO_ERR_INVALID, "forward_rate_extraction_selecting_step_gpu: null");
  hipLaunchKernelGGL(k_forward_extract, path_grid(paths, 256), dim3(256), 0, as_stream(stream),
                     all_forwards, idx, st, n_indices, paths, rates, out);
  return check_launch("forward_rate_extraction_selecting_step_gpu");
}

extern "C" int spot_measure_numeraires_extraction_gpu(float* all, float* some, int pathsInput,
                                                      int pathsOutput, int offsetForOutput,
                                                      int /*numberRates*/,
                                                      int numberExerciseTimes, int* exerciseIndices,
                                                      void* stream) {
  if (pathsInput < 0 || numberExerciseTimes < 0 || offsetForOutput < 0 ||
      offsetForOutput + pathsInput > pathsOutput)
    return set_error(KOO_ERR_INVALID, "spot_measure_numeraires_extraction_gpu: bad sizes");
  if (pathsInput == 0 || numberExerciseTimes == 0) return KOO_OK;
  if (!all || !some || !exerciseIndices)
    return set_error(KOO_ERR_INVALID, "spot_measure_numeraires_extraction_gpu: null pointer");
  hipLaunchKernelGGL(k_numeraire_extract, path_grid(pathsInput, 256), dim3(256), 0,
                     as_stream(stream), all, some, pathsInput, pathsOutput, offsetForOutput,
                     numberExerciseTimes, exerciseIndices);
  return check_launch("spot_measure_numeraires_extraction_gpu");
}
