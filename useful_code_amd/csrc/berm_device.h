// berm_device.h -- block bodies of the frozen-strategy Bermudan swaption
// price and pathwise adjoint (the forward with checkpoints and the reverse
// sweep), shared by the single-product and batched kernels and the
// persistent portfolio kernel.
#pragma once

#include "lmm_step.h"

namespace koo {
namespace {

using namespace lmm;

constexpr int kMaxBermDates = 40;
constexpr int kStatePrefetch = 8;

struct BermArgs {
  int E, itm_only, last;
  signed char e_of_step[kMaxRates];  // exercise index of each step, -1 if none
  float sh[kMaxBermDates][2], sc[kMaxBermDates][2];
};

__device__ __forceinline__ int wave_max_int(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
  return v;
}

// Floats of a product's checkpoint buffer before the step table that the
// single-product and batched forwards copy behind them.
template <int F>
__device__ __forceinline__ size_t berm_ckpt_floats(const Model& m) {
  const size_t rows = m.k == 1 ? m.N + 1 + 2 * F : m.N + 1;
  return static_cast<size_t>((m.N + m.k - 1) / m.k) * rows * static_cast<size_t>(m.P);
}

// Coefficients of the frozen strategy, date ei, function j, as FP32: from
// FP64 device memory (uniform addresses: scalar loads) or from an FP32 table.
struct CoefGlobal {
  const double* __restrict__ c;
  __device__ float operator()(int ei, int j) const { return static_cast<float>(c[6 * ei + j]); }
};
struct CoefTable {
  const float* t;
  __device__ float operator()(int ei, int j) const { return t[6 * ei + j]; }
};

// One block of the forward: paths [block * blockDim, ...) of product m.
// Agent: the block's price partial stored agent-coherent (persistent kernel).
template <int F, bool Agent = false, class Coef>
__device__ __forceinline__ void berm_fwd_block(const Model& m, const BermArgs& b, const Coef& cf,
                                               float* __restrict__ ckpt, float* __restrict__ lamt,
                                               int* __restrict__ tau_out,
                                               float* __restrict__ vout,
                                               double* __restrict__ part, int block) {
  extern __shared__ float lds[];
  __shared__ double red[4];
  const int bd = blockDim.x;
  const int p = block * bd + threadIdx.x;
  const bool valid = p < m.P;
  const int pc = valid ? p : m.P - 1;
  float* L = lds + threadIdx.x;
  float* ctab = lds + static_cast<size_t>(m.N) * bd;
  const size_t P = static_cast<size_t>(m.P);
  fill_step_table<F>(m, ctab);
  for (int r = 0; r < m.N; ++r) L[r * bd] = logf(m.f0[r] + m.disp[r]);
  __syncthreads();
  // Block 0 of a single-product or batched forward also copies the step
  // table behind the checkpoints, where the backward reads it (the
  // persistent kernel's backward keeps its own LDS copy: it runs in the same
  // launch as this forward).
  if (!Agent && ckpt && block == 0) {
    float* tab = ckpt + berm_ckpt_floats<F>(m);
    for (int i = threadIdx.x; i < step_table_floats(m.N); i += bd) tab[i] = ctab[i];
  }
  float invB = 1.0f, V = 0.0f;
  int tau = -1;
  bool done = !valid;
  const int rows = m.k == 1 ? m.N + 1 + 2 * F : m.N + 1;
  for (int s = 0; s <= b.last; ++s) {
    if (__all(done)) break;  // the wavefront's paths have all exercised
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
    const float fs = fast_exp(L[s * bd]) - m.disp[s];
    const int ei = b.e_of_step[s];
    if (ei >= 0 && !done) {
      float disc = 1.0f, ann = 0.0f, flt = 0.0f;
      for (int r = s; r < m.N; ++r) {
        const float fr = fast_exp(L[r * bd]) - m.disp[r];
        const float tr = m.taus[r];
        disc *= fast_rcp(1.0f + tr * fr);
        ann += tr * disc;
        flt += tr * fr * disc;
      }
      const float sr = flt / ann;
      const float ev = (sr - m.strike) * ann;
      const float x0 = (sr - b.sh[ei][0]) * b.sc[ei][0];
      const float x1 = (fs - b.sh[ei][1]) * b.sc[ei][1];
      const float cont = cf(ei, 0) + cf(ei, 1) * x0 + cf(ei, 2) * x1 + cf(ei, 3) * x0 * x0 +
                         cf(ei, 4) * x0 * x1 + cf(ei, 5) * x1 * x1;
      const float evp = ev > 0.0f ? ev : 0.0f;
      const bool allowed = !b.itm_only || evp > 0.0f;
      if (allowed && cont - evp < 0.0f) {
        done = true;
        tau = s;
        V = evp * invB;
        // payoff adjoint on the end-of-step log-rates r >= s
        float Pr = disc, Q = 0.0f;  // P_{r+1}, sum_{m>r} tau_{m-1} P_m
        for (int r = m.N - 1; r >= s; --r) {
          const float x = fast_exp(L[r * bd]);
          const float fr = x - m.disp[r];
          const float tr = m.taus[r];
          const float g = tr * fast_rcp(1.0f + tr * fr);
          Q += tr * Pr;
          lamt[static_cast<size_t>(r) * P + p] = ev > 0.0f ? g * (disc + m.strike * Q) * invB * x : 0.0f;
          Pr *= 1.0f + tr * fr;
        }
      }
    }
    invB *= fast_rcp(1.0f + m.taus[s] * fs);
  }
  if (valid) {
    tau_out[p] = tau;
    vout[p] = V;
  }
  block_sum_to<Agent>(valid ? static_cast<double>(V) : 0.0, red, part + block);
}

// One block of the backward (paths [block * blockDim, ...)).  Agent: the
// block's partials stored agent-coherent (persistent kernel).  LDS: the
// adjoints of rates 1..N-1 [N-1][lanes] (rate 0's in a register), the step
// table [N][N+1] in the persistent kernel only (the others read the
// forward's copy behind the checkpoints through the scalar cache), and one
// vega partial sum per wavefront and rate [wave][N].
template <int F, bool Agent = false>
__device__ __forceinline__ void berm_bwd_block(const Model& m, const float* __restrict__ vols,
                                               const float* __restrict__ ckpt,
                                               const float* __restrict__ lamt,
                                               const int* __restrict__ tau_in,
                                               const float* __restrict__ vin,
                                               float* __restrict__ scratch,
                                               double* __restrict__ part, int block) {
  extern __shared__ float lds[];
  __shared__ double red[4];
  const int bd = blockDim.x;
  const int nw = bd >> 6;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int p = block * bd + threadIdx.x;
  const bool valid = p < m.P;
  const int pc = valid ? p : m.P - 1;
  const int N = m.N;
  const size_t P = static_cast<size_t>(m.P);
  float lam0 = 0.0f;
  const LamReg0 lam{lds + threadIdx.x, bd, lam0};
  float* ltab = lds + static_cast<size_t>(N - 1) * bd;
  float* vslots = ltab + (Agent ? step_table_floats(N) : 0);
  float* myslots = vslots + static_cast<size_t>(wave) * N;
  if (Agent) fill_step_table<F>(m, ltab);
  for (int r = 1; r < N; ++r) lam.set(r, 0.0f);
  for (int i = threadIdx.x; i < nw * N; i += bd) vslots[i] = 0.0f;
  __syncthreads();
  const auto ctab = [&]() {
    if constexpr (Agent) return static_cast<const float*>(ltab);
    else return cmem(ckpt + berm_ckpt_floats<F>(m));
  }();

  const int tau = valid ? tau_in[p] : -1;
  const float V = valid ? vin[p] : 0.0f;
  // The wavefront's last exercise step, made a scalar (the butterfly leaves
  // it in every lane) so the loops below and their addresses are uniform.
  const int s_top = __builtin_amdgcn_readfirstlane(wave_max_int(tau));
  const size_t slot = static_cast<size_t>(N + 1 + 2 * F) * P;
  const bool every = m.k == 1;
  for (int seg = s_top < 0 ? -1 : s_top / m.k; seg >= 0; --seg) {
    const int s0 = seg * m.k;
    const int s1 = min(s_top + 1, s0 + m.k);
    const float* ck = ckpt + static_cast<size_t>(seg) * (every ? slot : (N + 1) * P) + pc;
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
        const float fsx = fast_exp(l_fix) - m.disp[s];
        dst[static_cast<size_t>(N) * P] =
            src[static_cast<size_t>(N) * P] * fast_rcp(1.0f + m.taus[s] * fsx);
      }
    }
    for (int s = s1 - 1; s >= s0; --s) {
      const int j = s - s0;
      const float* src = (j == 0) ? ck : scratch + j * slot + pc;
      const float* sums = (every ? ck : scratch + j * slot + pc) + static_cast<size_t>(N + 1) * P;
      if (s == tau)  // seed lambda with the payoff adjoint recorded at exercise
        for (int r = s; r < N; ++r) lam.set(r, lamt[static_cast<size_t>(r) * P + p]);
      float z[F], e[F], ep[F];
      load_z<F>(m, s, pc, z);
#pragma unroll
      for (int f = 0; f < F; ++f) {
        e[f] = sums[static_cast<size_t>(f) * P];
        ep[f] = sums[static_cast<size_t>(F + f) * P];
      }
      reverse_step<F, kStatePrefetch>(
          m, s, ctab + s * N, [&](int r) { return src[static_cast<size_t>(r) * P]; },
          lam, z, e, ep,
          [&](float& lp, float xs, float fs, float inv3) {
            (void)fs;
            if (s < tau) lp -= m.taus[s] * inv3 * V * xs;  // fixing inside B_tau
          },
          [&](int r, float w) {
            w = valid ? w : 0.0f;
            w += dpp_f32_quad_xor1(w);
            w += dpp_f32_quad_xor2(w);
            w += dpp_f32_row_half_mirror(w);
            w += dpp_f32_row_mirror(w);
            // every lane of a row holds the row's sum: the rows' last lanes
            // add them to the wavefront's slot (one LDS atomic, four lanes)
            if ((lane & 15) == 15) atomicAdd(myslots + r, w);
          });
    }
  }
  __syncthreads();
  for (int r = 0; r < N; ++r) {
    const float dl = lam.get(r) / (m.f0[r] + m.disp[r]);
    block_sum_to<Agent>(valid ? static_cast<double>(dl) : 0.0, red,
                        part + static_cast<size_t>(block) * 2 * N + r);
  }
  for (int r = threadIdx.x; r < N; r += bd) {
    double t = 0.0;
    for (int w = 0; w < nw; ++w) t += vslots[static_cast<size_t>(w) * N + r];
    double* o = part + static_cast<size_t>(block) * 2 * N + N + r;
    if (Agent)
      store_agent(o, t / vols[r]);
    else
      *o = t / vols[r];
  }
}

int fwd_threads(int N) { return N <= 50 ? 256 : (N <= 75 ? 128 : 64); }
int bwd_threads(int N) { return N <= 45 ? 256 : (N <= 71 ? 128 : 64); }
size_t fwd_lds(int N, int t) { return sizeof(float) * (static_cast<size_t>(N) * t + step_table_floats(N)); }
// lds_table: the persistent kernel's backward, which keeps the step table in
// LDS.
size_t bwd_lds(int N, int t, bool lds_table) {
  return sizeof(float) * (static_cast<size_t>(N - 1) * t + (lds_table ? step_table_floats(N) : 0) +
                          (t / 64) * N);
}

// Host checks and the strategy argument.
int check_berm(const char* who, int paths, int rates, int factors, int k) {
  if (paths < 1 || rates < 1 || rates > kMaxRates || factors < 1 || factors > kMaxFactors || k < 1)
    return set_error(KOO_ERR_INVALID,
                     "%s: need paths>=1, 1<=rates<=%d, 1<=factors<=%d, checkpoint_every>=1 "
                     "(got %d, %d, %d, %d)",
                     who, kMaxRates, kMaxFactors, paths, rates, factors, k);
  return KOO_OK;
}

// Host strategy -> kernel argument.
int make_berm_args(const char* who, int rates, int n_ex, const int* ex_steps,
                   const double* shift, const double* scale, int itm_only, BermArgs* b) {
  if (n_ex < 1 || n_ex > kMaxBermDates || !ex_steps)
    return set_error(KOO_ERR_INVALID, "%s: need 1<=exercise dates<=%d", who, kMaxBermDates);
  *b = BermArgs{};
  for (int s = 0; s < kMaxRates; ++s) b->e_of_step[s] = -1;
  for (int e = 0; e < n_ex; ++e) {
    const int s = ex_steps[e];
    if (s < 0 || s >= rates || (e > 0 && s <= ex_steps[e - 1]))
      return set_error(KOO_ERR_INVALID, "%s: exercise steps must increase within [0, %d)", who,
                       rates);
    b->e_of_step[s] = static_cast<signed char>(e);
    for (int i = 0; i < 2; ++i) {
      b->sh[e][i] = shift ? static_cast<float>(shift[2 * e + i]) : 0.0f;
      b->sc[e][i] = scale ? static_cast<float>(scale[2 * e + i]) : 1.0f;
    }
  }
  b->E = n_ex;
  b->itm_only = itm_only != 0;
  b->last = ex_steps[n_ex - 1];
  return KOO_OK;
}

}  // namespace
}  // namespace koo
