// lmm_price.hip -- fused LMM pricing with pathwise adjoint Greeks, gfx950.
//
// Forward (k_price_fwd): one lane per path evolves the whole rate curve with
// the predictor-corrector log-Euler step of LMM_evolver_all.cu:571-742 (spot
// measure, displaced diffusion), prices the product on the fly and never
// writes the rate cube: its only HBM traffic is the bridged variates (S,F,P)
// and, when asked, one checkpoint of the log-rate state every k steps.
//
// Backward (k_price_bwd): one reverse sweep per path.  For each checkpoint
// segment, last first, the k steps are recomputed from the checkpoint and
// each step's start state and final drift sums are kept in a scratch buffer;
// then the steps are reversed one by one, rates in descending order, with the
// running drift sums e_f, e^_f un-accumulated and the adjoint running sums
// E_f, E^_f accumulated in registers.  Deltas come from the adjoint of the
// initial log-rates; vegas from sum_{s,f} Abar(s,r,f) A(s,r,f) / sigma_r,
// whose per-rate contraction needs only quantities the reverse step already
// has (no per-factor Abar is formed).
//
// The model's uniform data (pseudo-roots, taus, displacements) is read with
// uniform addresses (scalar loads); the per-path state lives in LDS laid out
// [rate][lane] so wavefront accesses are conflict free; block partial sums
// are formed in FP64 and a final single pass sums them with compensation.
#include "lmm_step.h"

namespace koo {

namespace {

using namespace lmm;


// Caplet value at fixing, tau max(f-K,0)/(1+tau f), and its f-derivative.
__device__ __forceinline__ float caplet_value(float f, float tau, float K, float inv) {
  return f > K ? tau * (f - K) * inv : 0.0f;
}
__device__ __forceinline__ float caplet_deriv(float f, float tau, float K, float inv) {
  return f > K ? tau * (1.0f + tau * K) * inv * inv : 0.0f;
}

template <int F>
__global__ void __launch_bounds__(256)
    k_price_fwd(KOO_MODEL_PARAMS, float* __restrict__ ckpt, double* __restrict__ part) {
  KOO_MODEL_LOCAL;
  extern __shared__ float lds[];
  __shared__ double red[4];
  const int bd = blockDim.x;
  const int p = blockIdx.x * bd + threadIdx.x;
  const bool valid = p < m.P;
  const int pc = valid ? p : m.P - 1;
  float* L = lds + threadIdx.x;
  float* ctab = lds + static_cast<size_t>(m.N) * bd;
  const size_t P = static_cast<size_t>(m.P);
  fill_step_table<F>(m, ctab);
  for (int r = 0; r < m.N; ++r) L[r * bd] = logf(m.f0[r] + m.disp[r]);
  __syncthreads();
  float invB = 1.0f, y = 0.0f, yc = 0.0f;
  // With k == 1 every step is a checkpoint and also keeps its final drift
  // sums (rows N+1..N+2F), so the reverse sweep needs no recomputation.
  const int rows = m.k == 1 ? m.N + 1 + 2 * F : m.N + 1;
  // Block 0 also copies the step table behind the checkpoints, where the
  // reverse sweep reads it (the same bits the forward stepped with).
  if (ckpt && blockIdx.x == 0) {
    float* tab = ckpt + static_cast<size_t>((m.N + m.k - 1) / m.k) * rows * P;
    for (int i = threadIdx.x; i < step_table_floats(m.N); i += bd) tab[i] = ctab[i];
  }
  for (int s = 0; s < m.N; ++s) {
    const bool save = ckpt && valid && (s % m.k) == 0;
    float* c = save ? ckpt + static_cast<size_t>(s / m.k) * rows * P + p : nullptr;
    if (save) {
      for (int r = s; r < m.N; ++r) c[static_cast<size_t>(r) * P] = L[r * bd];
      c[static_cast<size_t>(m.N) * P] = invB;
    }
    float z[F], e[F], ep[F];
    load_z<F>(m, s, pc, z);
    pc_step<F>(m, s, ctab + s * m.N, z, e, ep, [&](int r) { return L[r * bd]; },
               [&](int r, float v) { L[r * bd] = v; });
    if (save && m.k == 1) {
#pragma unroll
      for (int f = 0; f < F; ++f) {
        c[static_cast<size_t>(m.N + 1 + f) * P] = e[f];
        c[static_cast<size_t>(m.N + 1 + F + f) * P] = ep[f];
      }
    }
    const float tau = m.taus[s];
    const float fs = fast_exp(L[s * bd]) - m.disp[s];
    const float inv = fast_rcp(1.0f + tau * fs);
    // Kahan-compensated path sum of the deflated caplet values.
    const float term = caplet_value(fs, tau, m.strike, inv) * invB - yc;
    const float t = y + term;
    yc = (t - y) - term;
    y = t;
    invB *= inv;
  }
  block_sum_to(valid ? static_cast<double>(y) : 0.0, red, part + blockIdx.x);
}

// ----------------------------------------------------------------------
// Backward.  scratch slot j (j < k) holds rows [0,N): start log-rates of step
// s0+j, row N: 1/B at that step, rows N+1..N+2F: final sums e, e^ of step
// s0+j.  Slot 0's state rows are not used (the checkpoint is read instead).
// The step table (fixed drifts [N][N], then kr [N]) is the forward's copy
// behind the checkpoints, read with scalar loads.  LDS: adjoints of rates
// 1..N-1 [N-1][lanes] (rate 0 only changes in step 0 and lives in a
// register), one vega partial sum per wavefront and rate [wave][N] (each
// wavefront folds its lanes to the four 16-lane row sums with DPP, and the
// last lane of each row adds its row's sum to the slot: one LDS atomic with
// four lanes on one address; repeated runs give identical sums).  At 40 rates and 256 lanes
// that is 40 608 bytes with the block-sum words: four blocks per CU.
// ----------------------------------------------------------------------

// Depth of the queue of start log-rates read from checkpoints or scratch in
// the backward kernel (HBM latency over the per-rate work of the resident
// wavefronts).  Measured at 2^20 paths, 40 rates, 5 factors with three
// 256-lane blocks per CU: 4.39 ms at 4, 3.97 at 8, 3.89 at 12 and 16; with
// four per CU, 8, 12 and 16 are equal (3.70-3.79 ms).
constexpr int kStatePrefetch = 12;

template <int F>
__global__ void __launch_bounds__(256)
    k_price_bwd(KOO_MODEL_PARAMS, const float* __restrict__ vols, const float* __restrict__ ckpt,
                const float* __restrict__ ctab, float* __restrict__ scratch,
                double* __restrict__ part) {
  KOO_MODEL_LOCAL;
  extern __shared__ float lds[];
  __shared__ double red[4];
  const int bd = blockDim.x;
  const int nw = bd >> 6;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int p = blockIdx.x * bd + threadIdx.x;
  const bool valid = p < m.P;
  const int pc = valid ? p : m.P - 1;
  const int N = m.N;
  const size_t P = static_cast<size_t>(m.P);
  float lam0 = 0.0f;
  const LamReg0 lam{lds + threadIdx.x, bd, lam0};             // rates 1..N-1: [N-1][bd]
  float* vslots = lds + static_cast<size_t>(N - 1) * bd;      // [nw][N]
  float* myslots = vslots + static_cast<size_t>(wave) * N;
  for (int r = 1; r < N; ++r) lam.set(r, 0.0f);
  for (int i = threadIdx.x; i < nw * N; i += bd) vslots[i] = 0.0f;
  __syncthreads();

  const size_t slot = static_cast<size_t>(N + 1 + 2 * F) * P;
  const int nseg = (N + m.k - 1) / m.k;
  float tail = 0.0f;
  for (int seg = nseg - 1; seg >= 0; --seg) {
    const int s0 = seg * m.k;
    const int s1 = min(N, s0 + m.k);
    const bool every = m.k == 1;  // checkpoint holds the sums: no recompute
    const float* ck = ckpt + static_cast<size_t>(seg) * (every ? slot : (N + 1) * P) + pc;
    // Recompute the segment: state (and 1/B) at the start of step s0+j+1
    // into slot j+1, final drift sums of step s0+j into slot j.
    for (int s = s0; s < (every ? s0 : s1); ++s) {
      const int j = s - s0;
      const float* src = (j == 0) ? ck : scratch + j * slot + pc;
      float* dst = scratch + (j + 1) * slot + p;
      float* sums = scratch + j * slot + static_cast<size_t>(N + 1) * P + p;
      const bool keep = valid && (s + 1 < s1);
      float z[F], e[F], ep[F];
      float l_fix = 0.0f;
      load_z<F>(m, s, pc, z);
      pc_step<F, kStatePrefetch>(m, s, ctab + s * N, z, e, ep,
                 [&](int r) { return src[static_cast<size_t>(r) * P]; },
                 [&](int r, float v) {
                   if (r == s) l_fix = v;
                   else if (keep) dst[static_cast<size_t>(r) * P] = v;
                 });
      if (valid) {
#pragma unroll
        for (int f = 0; f < F; ++f) {
          sums[static_cast<size_t>(f) * P] = e[f];
          sums[static_cast<size_t>(F + f) * P] = ep[f];
        }
      }
      if (keep) {
        const float fs = fast_exp(l_fix) - m.disp[s];
        dst[static_cast<size_t>(N) * P] =
            src[static_cast<size_t>(N) * P] * fast_rcp(1.0f + m.taus[s] * fs);
      }
    }
    // Reverse the segment's steps.
    for (int s = s1 - 1; s >= s0; --s) {
      const int j = s - s0;
      const float* src = (j == 0) ? ck : scratch + j * slot + pc;
      const float* sums = (every ? ck : scratch + j * slot + pc) + static_cast<size_t>(N + 1) * P;
      const float* cs = ctab + s * N;
      const float invBs = src[static_cast<size_t>(N) * P];
      float z[F], e[F], ep[F];
      load_z<F>(m, s, pc, z);
#pragma unroll
      for (int f = 0; f < F; ++f) {
        e[f] = sums[static_cast<size_t>(f) * P];
        ep[f] = sums[static_cast<size_t>(F + f) * P];
      }
      reverse_step<F, kStatePrefetch>(
          m, s, cs, [&](int r) { return src[static_cast<size_t>(r) * P]; }, lam, z, e, ep,
          [&](float& lp, float xs, float fs, float inv3) {
            // The rate fixing at the end of this step: caplet value and the
            // numeraire it enters for later fixings.
            const float tau = m.taus[s];
            const float v = caplet_value(fs, tau, m.strike, inv3);
            const float dv = caplet_deriv(fs, tau, m.strike, inv3);
            lp += (dv * invBs - tau * inv3 * tail) * xs;
            tail += v * invBs;
          },
          [&](int r, float w) {
            w = valid ? w : 0.0f;
            w += dpp_f32_quad_xor1(w);
            w += dpp_f32_quad_xor2(w);
            w += dpp_f32_row_half_mirror(w);
            w += dpp_f32_row_mirror(w);
            if ((lane & 15) == 15) atomicAdd(myslots + r, w);
          });
    }
  }
  __syncthreads();
  // Deltas: d/df0_r = lam_r / (f0_r + d_r); vegas: slot sums / sigma_r.
  for (int r = 0; r < N; ++r) {
    const float dl = lam.get(r) / (m.f0[r] + m.disp[r]);
    block_sum_to(valid ? static_cast<double>(dl) : 0.0, red,
                 part + static_cast<size_t>(blockIdx.x) * 2 * N + r);
  }
  for (int r = threadIdx.x; r < N; r += bd) {
    double t = 0.0;
    for (int w = 0; w < nw; ++w) t += vslots[static_cast<size_t>(w) * N + r];
    part[static_cast<size_t>(blockIdx.x) * 2 * N + N + r] = t / vols[r];
  }
}

// Lanes per block keeping the dynamic LDS within 60 KiB:
// forward N*lanes + N*(N+1) floats, backward (N-1)*lanes + (lanes/64)*N
// (256 lanes up to 60 rates; 58 keeps a margin).
int fwd_threads(int N) { return N <= 50 ? 256 : (N <= 75 ? 128 : 64); }
int bwd_threads(int N) { return N <= 58 ? 256 : 128; }
size_t fwd_lds(int N, int t) { return sizeof(float) * (static_cast<size_t>(N) * t + step_table_floats(N)); }
size_t bwd_lds(int N, int t) {
  return sizeof(float) * (static_cast<size_t>(N - 1) * t + (t / 64) * N);
}
// Floats of the checkpoint buffer before the step table the forward copies
// behind them.
size_t ckpt_floats(int paths, int N, int F, int k) {
  const size_t rows = k == 1 ? N + 1 + 2 * F : N + 1;
  return static_cast<size_t>((N + k - 1) / k) * rows * paths;
}

int check_model(const char* who, int paths, int rates, int factors, int k, int product) {
  if (paths < 1 || rates < 1 || rates > kMaxRates || factors < 1 || factors > kMaxFactors || k < 1)
    return set_error(KOO_ERR_INVALID,
                     "%s: need paths>=1, 1<=rates<=%d, 1<=factors<=%d, checkpoint_every>=1 "
                     "(got %d, %d, %d, %d)",
                     who, kMaxRates, kMaxFactors, paths, rates, factors, k);
  if (product != KOO_PRODUCT_CAPLETS)
    return set_error(KOO_ERR_INVALID, "%s: unknown product %d", who, product);
  return KOO_OK;
}

}  // namespace
}  // namespace koo

using namespace koo;

extern "C" int lmm_price_workspace(int paths, int rates, int factors, int checkpoint_every,
                                   size_t* checkpoint_bytes, size_t* scratch_bytes,
                                   size_t* partial_bytes) {
  if (int e = check_model("lmm_price_workspace", paths, rates, factors, checkpoint_every,
                          KOO_PRODUCT_CAPLETS))
    return e;
  const size_t P = static_cast<size_t>(paths);
  const int k = checkpoint_every;
  if (checkpoint_bytes)
    *checkpoint_bytes =
        sizeof(float) * (ckpt_floats(paths, rates, factors, k) + step_table_floats(rates));
  // k == 1 needs no recompute scratch; one float keeps the pointer non-null.
  if (scratch_bytes)
    *scratch_bytes = k == 1 ? sizeof(float)
                            : sizeof(float) * static_cast<size_t>(k < rates ? k : rates) *
                                  (rates + 1 + 2 * factors) * P;
  const size_t bf = (P + fwd_threads(rates) - 1) / fwd_threads(rates);
  const size_t bb = (P + bwd_threads(rates) - 1) / bwd_threads(rates);
  const size_t need = bf > bb * 2 * rates ? bf : bb * 2 * rates;
  if (partial_bytes) *partial_bytes = sizeof(double) * need;
  return KOO_OK;
}

extern "C" int lmm_price_fwd(int paths, int rates, int factors, int checkpoint_every,
                             const float* variates, const float* initial_rates, const float* taus,
                             const float* displacements, const float* pseudo_roots, int product,
                             const float* product_aux, float* checkpoints, double* partials,
                             double* price_out, void* stream) {
  if (int e = check_model("lmm_price_fwd", paths, rates, factors, checkpoint_every, product))
    return e;
  if (!variates || !initial_rates || !taus || !displacements || !pseudo_roots || !product_aux ||
      !partials || !price_out)
    return set_error(KOO_ERR_INVALID, "lmm_price_fwd: null pointer");
  const float strike = product_aux[0];
  Model m{variates, cmem(initial_rates), cmem(taus), cmem(displacements), cmem(pseudo_roots),
          strike, paths, rates,
          checkpoint_every};
  const int threads = fwd_threads(rates);
  const int blocks = (paths + threads - 1) / threads;
  const size_t lds = fwd_lds(rates, threads);
  hipStream_t st = as_stream(stream);
  switch (factors) {
#define KOO_CASE(F)                                                                              \
  case F:                                                                                        \
    hipLaunchKernelGGL(k_price_fwd<F>, dim3(blocks), dim3(threads), lds, st, m.z, m.f0, m.taus, \
                       m.disp, m.A, m.strike, m.P, m.N, m.k, checkpoints, partials);            \
    break;
    KOO_CASE(1) KOO_CASE(2) KOO_CASE(3) KOO_CASE(4) KOO_CASE(5) KOO_CASE(6) KOO_CASE(7)
    KOO_CASE(8)
#undef KOO_CASE
  }
  hipLaunchKernelGGL(k_finalize_cols, dim3(1), dim3(256), 0, st, partials, blocks, 1,
                     1.0 / static_cast<double>(paths), price_out);
  return check_launch("lmm_price_fwd");
}

extern "C" int lmm_price_bwd(int paths, int rates, int factors, int checkpoint_every,
                             const float* variates, const float* initial_rates, const float* taus,
                             const float* displacements, const float* pseudo_roots,
                             const float* vols, int product, const float* product_aux,
                             const float* checkpoints, float* segment_scratch, double* partials,
                             double* greeks_out, void* stream) {
  if (int e = check_model("lmm_price_bwd", paths, rates, factors, checkpoint_every, product))
    return e;
  if (!variates || !initial_rates || !taus || !displacements || !pseudo_roots || !vols ||
      !product_aux || !checkpoints || !segment_scratch || !partials || !greeks_out)
    return set_error(KOO_ERR_INVALID, "lmm_price_bwd: null pointer");
  const float strike = product_aux[0];
  Model m{variates, cmem(initial_rates), cmem(taus), cmem(displacements), cmem(pseudo_roots),
          strike, paths, rates,
          checkpoint_every};
  const int threads = bwd_threads(rates);
  const int blocks = (paths + threads - 1) / threads;
  const size_t lds = bwd_lds(rates, threads);
  const float* tab = checkpoints + ckpt_floats(paths, rates, factors, checkpoint_every);
  hipStream_t st = as_stream(stream);
  switch (factors) {
#define KOO_CASE(F)                                                                              \
  case F:                                                                                        \
    hipLaunchKernelGGL(k_price_bwd<F>, dim3(blocks), dim3(threads), lds, st, m.z, m.f0, m.taus, \
                       m.disp, m.A, m.strike, m.P, m.N, m.k, vols, checkpoints, tab,            \
                       segment_scratch, partials);                                              \
    break;
    KOO_CASE(1) KOO_CASE(2) KOO_CASE(3) KOO_CASE(4) KOO_CASE(5) KOO_CASE(6) KOO_CASE(7)
    KOO_CASE(8)
#undef KOO_CASE
  }
  hipLaunchKernelGGL(k_finalize_cols, dim3(2 * rates), dim3(256), 0, st, partials, blocks,
                     2 * rates, 1.0 / static_cast<double>(paths), greeks_out);
  return check_launch("lmm_price_bwd");
}
