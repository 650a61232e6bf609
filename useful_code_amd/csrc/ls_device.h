// ls_device.h -- device pieces of the Longstaff-Schwartz passes shared by
// the LS kernels and the persistent portfolio kernel: the Bermudan payer
// swaption path quantities of one block, and one block's work in one round
// of the first pass (the later date's decision and deflation, this date's
// normal equations on MFMA, or the final decision and price partial).
#pragma once

#include "ls_common.h"

namespace koo {
namespace {

using namespace lmm;

// ---------------------------------------------------------------------------
// Bermudan payer swaption path quantities.
// ---------------------------------------------------------------------------
struct ExMap {
  signed char e_of_step[kMaxRates];  // exercise index of each step, -1 if none
  int last;                          // last exercise step
};

template <int F>
__device__ __forceinline__ void berm_paths_block(const Model& m, const ExMap& ex,
                                                 float* __restrict__ bv, float* __restrict__ evo,
                                                 float* __restrict__ numo, int block) {
  extern __shared__ float lds[];
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
  float B = 1.0f;
  for (int s = 0; s <= ex.last; ++s) {
    float z[F], e[F], ep[F];
    load_z<F>(m, s, pc, z);
    pc_step<F>(m, s, ctab + s * m.N, z, e, ep, [&](int r) { return L[r * bd]; },
               [&](int r, float v) { L[r * bd] = v; });
    const float tau = m.taus[s];
    const float fs = fast_exp(L[s * bd]) - m.disp[s];
    const int ei = ex.e_of_step[s];
    if (ei >= 0) {
      // P_{r+1} = P_r / (1 + tau_r f_r) from P_s = 1; annuity sum tau_r P_{r+1};
      // swap rate as the annuity-weighted forward average, which equals
      // (1 - P_N) / annuity without its cancellation.
      float disc = 1.0f, ann = 0.0f, flt = 0.0f;
      for (int r = s; r < m.N; ++r) {
        const float fr = fast_exp(L[r * bd]) - m.disp[r];
        const float tr = m.taus[r];
        disc *= fast_rcp(1.0f + tr * fr);
        ann += tr * disc;
        flt += tr * fr * disc;
      }
      const float sr = flt / ann;
      const float v = (sr - m.strike) * ann;
      if (valid) {
        const size_t o = static_cast<size_t>(ei) * P + p;
        bv[static_cast<size_t>(2 * ei) * P + p] = sr;
        bv[static_cast<size_t>(2 * ei + 1) * P + p] = fs;
        evo[o] = v > 0.0f ? v : 0.0f;
        numo[o] = B;
      }
    }
    B *= 1.0f + tau * fs;
  }
}

// ---------------------------------------------------------------------------
// LS first pass.  Round r = 0..E handles exercise date e = E - 1 - r (round
// E: date 0's exercise decision and the price).  Every block of a round first
// sums the previous round's block partials and solves that date's normal
// equations itself (the same partials in the same order, so the same
// coefficients in every block; block 0 stores them), then does its paths'
// work for date e -- the later date's exercise decision and deflation of the
// running values, this date's normal equations on MFMA -- into its own column
// of the round's partials.  Partials: the upper triangle of the Gram, FP64,
// in three slots used in turn (round r writes slot r % 3).  Either one
// launch per round or one persistent launch; the per-block arithmetic is the
// same, so the results are too.
// ---------------------------------------------------------------------------
struct DateArgs {
  int e, E, P, max_vars, maxd;
  int v, k;    // variables and basis functions at date e (0 when e < 0)
  int vn, kn;  // at date e + 1 (when e + 1 < E)
  float sh[kMaxBasisVars], sc[kMaxBasisVars];    // x~ = (x - sh) * sc at date e
  float shn[kMaxBasisVars], scn[kMaxBasisVars];  // at date e + 1
  int cross, itm_only;
};

// Continuation value at a date for one path: basis of the normalised
// variables dotted with that date's coefficients (FP64 in HBM, FP32 here).
// The same from the date's variables already loaded (raw, before
// normalisation; entries from v on are ignored).
__device__ __forceinline__ float continuation_raw(const float (&xr)[kMaxBasisVars], int v, int k,
                                                  bool cross, const float* sh, const float* sc,
                                                  const double* __restrict__ beta) {
  float x[kMaxBasisVars] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int i = 0; i < kMaxBasisVars; ++i)
    if (i < v) x[i] = (xr[i] - sh[i]) * sc[i];
  float c = 0.0f;
  for (int j = 0; j < k; ++j)
    c += basis_value(j, v, cross, x[0], x[1], x[2], x[3]) * static_cast<float>(beta[j]);
  return c;
}

__device__ __forceinline__ float continuation(const float* __restrict__ bv, size_t row0, size_t P,
                                              int p, int v, int k, bool cross, const float* sh,
                                              const float* sc, const double* __restrict__ beta) {
  float xr[kMaxBasisVars] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int i = 0; i < kMaxBasisVars; ++i)
    if (i < v) xr[i] = bv[(row0 + i) * P + p];
  return continuation_raw(xr, v, k, cross, sh, sc, beta);
}

// The same from variables already loaded (raw, before normalisation) and
// coefficients already in registers (bf, FP32); KN >= k.
template <int KN>
__device__ __forceinline__ float continuation_of(const float (&xr)[kMaxBasisVars], int v, int k,
                                                 bool cross, const float* sh, const float* sc,
                                                 const float (&bf)[KN]) {
  float x[kMaxBasisVars] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int i = 0; i < kMaxBasisVars; ++i)
    if (i < v) x[i] = (xr[i] - sh[i]) * sc[i];
  float c = 0.0f;
#pragma unroll
  for (int j = 0; j < KN; ++j)
    if (j < k) c += basis_value(j, v, cross, x[0], x[1], x[2], x[3]) * bf[j];
  return c;
}

// One path's inputs to a round, loaded ahead of their use.
struct PathIn {
  float y;                      // running value (when date e + 1 exists)
  float xn[kMaxBasisVars];      // date e + 1: raw variables,
  float evn, numn, cashn;       //   exercise value, numeraire, cash flow
  float x[kMaxBasisVars];       // date e (when e >= 0)
  float eve, nume, cashe;
};

// Issues the loads of path p's inputs to the round of `a` (p < P); the
// running value from Y unless Y is null (kept in a register by the caller).
__device__ __forceinline__ PathIn load_path(const DateArgs& a, const float* __restrict__ bv,
                                            const float* __restrict__ cash,
                                            const float* __restrict__ ev,
                                            const float* __restrict__ num,
                                            const float* __restrict__ Y, int p) {
  PathIn in{};
  const size_t P = static_cast<size_t>(a.P);
  if (a.e + 1 < a.E) {
    const size_t on = static_cast<size_t>(a.e + 1) * P + p;
    const size_t rn = static_cast<size_t>(a.e + 1) * a.max_vars;
    if (Y) in.y = Y[p];
#pragma unroll
    for (int i = 0; i < kMaxBasisVars; ++i)
      if (i < a.vn) in.xn[i] = bv[(rn + i) * P + p];
    in.evn = ev[on];
    in.numn = num[on];
    if (cash) in.cashn = cash[on];
  }
  if (a.e >= 0) {
    const size_t oe = static_cast<size_t>(a.e) * P + p;
    const size_t re = static_cast<size_t>(a.e) * a.max_vars;
#pragma unroll
    for (int i = 0; i < kMaxBasisVars; ++i)
      if (i < a.v) in.x[i] = bv[(re + i) * P + p];
    in.eve = ev[oe];
    in.nume = num[oe];
    if (cash) in.cashe = cash[oe];
  }
  return in;
}

// The later date's exercise decision on a path under its coefficients beta_n
// (updateDeflatedNextStepValuesMulti_gold, gold/MultiLS_regression.h:46-115):
// the exercise value if the strategy exercises, else the running value plus
// the later date's cash flow.
template <int KN>
__device__ __forceinline__ float later_value(const DateArgs& a, const PathIn& in, bool cash,
                                             const float (&bf)[KN]) {
  const float cont = continuation_of<KN>(in.xn, a.vn, a.kn, a.cross != 0, a.shn, a.scn, bf);
  const bool allowed = !a.itm_only || in.evn > 0.0f;
  if (allowed && in.evn > cont) return in.evn;
  return cash ? in.y + in.cashn : in.y;
}

// The later date's coefficients (LDS, FP64) as FP32 registers, KN >= kn.
template <int KN>
__device__ __forceinline__ void coeffs_f32(const double* beta, int kn, float (&bf)[KN]) {
#pragma unroll
  for (int j = 0; j < KN; ++j) bf[j] = j < kn ? static_cast<float>(beta[j]) : 0.0f;
}

// The path work of one block (of nblocks) for date e >= 0 given the later
// date's coefficients in beta_n (LDS; unused at the last date): the later
// date's decision and deflation of the running values (to date e,
// updateDeflationOfContinuationValues_gold, gold/LS_regression.cpp:98-115),
// then this date's normal equations on MFMA into the block's column of out.
// The block's paths are walked 64 per wavefront, the next chunk's loads
// issued before the current chunk is used; `first` is the first chunk,
// loaded by the caller.  yin: the running value of the lane's only chunk
// is kept in the register yr instead of Y.  Each wavefront stages its 64
// paths' weighted variables and target in its own 2 KB of LDS (basis_factors'
// vector), from which the lane of MFMA row j reads the two factors of
// function j for the path it supplies (lds_d: 256 doubles of LDS per
// wavefront, reused for the block's Gram).  Agent: agent-coherent
// partial stores (the persistent kernels).  KN >= the later date's basis size.
template <int GROUPS, bool Agent, int KN>
__device__ __forceinline__ void ls_date_work(const DateArgs& a, const float* __restrict__ bv,
                                             const float* __restrict__ cash,
                                             const float* __restrict__ ev,
                                             const float* __restrict__ num,
                                             const double* beta_n, float* __restrict__ Y,
                                             bool yin, float& yr, double* __restrict__ out,
                                             int block, int nblocks, const PathIn& first,
                                             double* lds_d /* LDS, 256 per wavefront */) {
  constexpr int kPathsPerOp = 4 * GROUPS;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const bool cross = a.cross != 0, itm = a.itm_only != 0;
  const bool later = a.e + 1 < a.E;
  const int n = a.k + 1;  // basis functions + target
  const int j = GROUPS == 2 ? (lane & 7) : (lane & 15);
  const int sub = GROUPS == 2 ? ((lane >> 3) & 1) * 4 + (lane >> 4) : (lane >> 4);
  float bf[KN];
  coeffs_f32<KN>(beta_n, later ? a.kn : 0, bf);
  // per path: basis_factors' vector (8 floats, 32 bytes)
  float* stage = reinterpret_cast<float*>(lds_d + wave * 256);
  int ia, ib;
  basis_factors(j, a.v, a.k, cross, ia, ib);
  // each MFMA's 4-path sums (FP32) accumulated in FP64
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  const int stride = nblocks * nw * 64;
  PathIn cur = first;
  if (yin) cur.y = yr;
  for (int base = (block * nw + wave) * 64; base < a.P; base += stride) {
    const int nbase = base + stride;
    PathIn nxt{};
    if (nbase < a.P)  // uniform per wavefront
      nxt = load_path(a, bv, cash, ev, num, Y, min(nbase + lane, a.P - 1));
    const int p = base + lane;
    float x0 = 0.0f, x1 = 0.0f, x2 = 0.0f, x3 = 0.0f, t = 0.0f, w = 0.0f;
    if (p < a.P) {
      float y = 0.0f;
      if (later) y = later_value<KN>(a, cur, cash != nullptr, bf) * (cur.nume / cur.numn);
      if (yin)
        yr = y;
      else
        Y[p] = y;
      w = (!itm || cur.eve > 0.0f) ? 1.0f : 0.0f;
      t = cur.cashe + y;
      if (a.v > 0) x0 = (cur.x[0] - a.sh[0]) * a.sc[0];
      if (a.v > 1) x1 = (cur.x[1] - a.sh[1]) * a.sc[1];
      if (a.v > 2) x2 = (cur.x[2] - a.sh[2]) * a.sc[2];
      if (a.v > 3) x3 = (cur.x[3] - a.sh[3]) * a.sc[3];
    }
    reinterpret_cast<floatx4*>(stage)[2 * lane] = floatx4{w, w * x0, w * x1, w * x2};
    reinterpret_cast<floatx4*>(stage)[2 * lane + 1] = floatx4{w * x3, w * t, 0.0f, 1.0f};
    // 64 paths of this wave -> 64 / (4 GROUPS) MFMAs; lane l takes row j of path `sub`.
#pragma unroll 4
    for (int q = 0; q < 64 / kPathsPerOp; ++q) {
      const float* sp = stage + 8 * (q * kPathsPerOp + sub);
      const float v = sp[ia] * sp[ib];
      const floatx4 d = mfma16x16x4(v, v, floatx4{0.0f, 0.0f, 0.0f, 0.0f});
#pragma unroll
      for (int i = 0; i < 4; ++i) acc[i] += static_cast<double>(d[i]);
    }
    cur = nxt;
  }
  block_gram_tri<GROUPS, Agent>(acc, n, lds_d, out, block, nblocks);
}

// Round E of one block: date 0's decision under its coefficients beta0 (LDS),
// then sum_p Y_p / B_0 into the block's price partial out[block].  `a` is the
// round's arguments (e = -1, date 0 as the later date); yin, yr as for
// ls_date_work.  Agent as for ls_date_work.
template <bool Agent, int KN>
__device__ __forceinline__ void ls_final_work(const DateArgs& a, const float* __restrict__ bv,
                                              const float* __restrict__ cash,
                                              const float* __restrict__ ev,
                                              const float* __restrict__ num, const double* beta0,
                                              const float* __restrict__ Y, bool yin, float yr,
                                              double* __restrict__ out, int block, int nblocks,
                                              const PathIn& first) {
  __shared__ double red[kGramThreads / 64];
  float bf[KN];
  coeffs_f32<KN>(beta0, a.kn, bf);
  double s = 0.0;
  const int stride = nblocks * blockDim.x;
  PathIn cur = first;
  if (yin) cur.y = yr;
  for (int base = block * blockDim.x + (threadIdx.x & ~63); base < a.P; base += stride) {
    const int nbase = base + stride;
    PathIn nxt{};
    if (nbase < a.P)
      nxt = load_path(a, bv, cash, ev, num, Y, min(nbase + (threadIdx.x & 63), a.P - 1));
    if (base + static_cast<int>(threadIdx.x & 63) < a.P)
      s += static_cast<double>(later_value<KN>(a, cur, cash != nullptr, bf)) /
           static_cast<double>(cur.numn);
    cur = nxt;
  }
  block_sum_to<Agent>(s, red, out + block);
}

// Partial slot of round r in a buffer of three, each kTriMax entries x nb.
__host__ __device__ inline size_t ls_slot(int r, int nb) {
  return static_cast<size_t>(r % 3) * kTriMax * nb;
}

}  // namespace
}  // namespace koo
