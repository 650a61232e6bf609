// lmm_step.h -- device building blocks shared by the fused LMM kernels
// (caplet price/adjoint and the Bermudan swaption paths): wavefront and block
// reductions, the model arguments, and the predictor-corrector log-Euler step
// of LMM_evolver_all.cu:571-742 (spot measure, displaced diffusion).
#pragma once

#include "koo_common.h"

namespace koo {
namespace lmm {
// Internal linkage: every kernel source gets its own copy.
namespace {

constexpr int kMaxRates = 90;
constexpr int kMaxFactors = 8;

__device__ __forceinline__ float fast_exp(float x) { return __expf(x); }
__device__ __forceinline__ float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }

// Drift coefficient g = tau (f + d) / (1 + tau f) of a displaced rate from
// its log l = log(f + d): g = 1 / (1 + kr exp(-l)) with kr = 1/tau - d (one
// multiply-add between the two hardware transcendentals).  Its derivative in
// l is g (1 - g).
__device__ __forceinline__ float drift_coeff(float l, float kr) {
  return fast_rcp(1.0f + kr * fast_exp(-l));
}

__device__ __forceinline__ float dpp_f32_quad_xor1(float x) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0xB1, 0xF, 0xF, false));
}
__device__ __forceinline__ float dpp_f32_quad_xor2(float x) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x4E, 0xF, 0xF, false));
}
__device__ __forceinline__ float dpp_f32_row_half_mirror(float x) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x141, 0xF, 0xF, false));
}
__device__ __forceinline__ float dpp_f32_row_mirror(float x) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x140, 0xF, 0xF, false));
}
__device__ __forceinline__ float readlane_f32(float x, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), lane));
}

// Sum over the 64 lanes of a wavefront with DPP row operations (quad
// butterflies, then half-row and row mirrors give each 16-lane row its sum;
// the four row sums are read from lanes 0, 16, 32, 48).  Returns the total,
// uniform across the wavefront.  All 64 lanes must be active.
__device__ __forceinline__ float wave_sum_f32(float v) {
  v += dpp_f32_quad_xor1(v);
  v += dpp_f32_quad_xor2(v);
  v += dpp_f32_row_half_mirror(v);
  v += dpp_f32_row_mirror(v);
  return (readlane_f32(v, 0) + readlane_f32(v, 16)) + (readlane_f32(v, 32) + readlane_f32(v, 48));
}

__device__ __forceinline__ double wave_sum_f64(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Agent-scope (device-wide) coherent store and load of a double: the store
// writes through the XCD's L2, the load misses it, so values exchanged
// between blocks of one running kernel need no cache write-back or
// invalidation (used by the persistent kernels for their block partials).
__device__ __forceinline__ void store_agent(double* p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double load_agent(const double* p) {
  return __hip_atomic_load(const_cast<double*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Block-wide FP64 sum of one value per lane into part[col]; every lane of the
// block must call it (contains barriers).  Agent: the store is agent-coherent.
template <bool Agent = false>
__device__ __forceinline__ void block_sum_to(double v, double* scratch_lds, double* out) {
  v = wave_sum_f64(v);
  const int w = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) scratch_lds[w] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int i = 0; i < static_cast<int>(blockDim.x >> 6); ++i) t += scratch_lds[i];
    if (Agent)
      store_agent(out, t);
    else
      *out = t;
  }
}

// Constant address space (4): the model's curve and pseudo-roots are read
// through it, since no kernel writes them, so that their uniform reads use
// the scalar cache also where the model comes from a device descriptor (the
// batched and persistent kernels), not only from __restrict__ arguments.
#if defined(__HIP_DEVICE_COMPILE__)
#define KOO_CONSTANT __attribute__((address_space(4)))
#else
#define KOO_CONSTANT
#endif
typedef const float KOO_CONSTANT* c_cfloat;
template <class P>
__host__ __device__ __forceinline__ c_cfloat cmem(P p) {
  return (c_cfloat)p;
}

struct Model {
  const float* z;     // (S,F,P) bridged variates
  c_cfloat f0;        // N
  c_cfloat taus;      // N
  c_cfloat disp;      // N
  c_cfloat A;         // (S,N,F)
  float strike;
  int P, N, k;
};

// The step table in LDS (read back with uniform addresses): rows 0..N-1 hold
// the fixed drifts c(s,r) = -0.5 sum_f A(s,r,f)^2, row N the drift
// coefficients' constants kr = 1/tau_r - d_r.  The steps are given row s and
// find row N at N - s rows past it.
__host__ __device__ constexpr size_t step_table_floats(int N) {
  return static_cast<size_t>(N) * N + N;
}

// Entry i of the step table (one function for every copy of the table, so
// an LDS copy and a global copy hold the same bits).
template <int F>
__device__ __forceinline__ float step_table_entry(const Model& m, int i) {
  const int nn = m.N * m.N;
  if (i >= nn) {
    const int r = i - nn;
    return 1.0f / m.taus[r] - m.disp[r];
  }
  c_cfloat a = m.A + static_cast<size_t>(i) * F;
  float c = 0.0f;
#pragma unroll
  for (int f = 0; f < F; ++f) c -= 0.5f * a[f] * a[f];
  return c;
}

// Every lane of the block takes part; the caller then synchronises.
template <int F>
__device__ __forceinline__ void fill_step_table(const Model& m, float* ctab) {
  for (int i = threadIdx.x; i < m.N * m.N + m.N; i += blockDim.x)
    ctab[i] = step_table_entry<F>(m, i);
}

// Log-rate adjoints: rate 0 in a register (it only changes in step 0), rate
// r >= 1 at p[(r - 1) * bd] in LDS (lane-minor rows).  The rate index is
// uniform, so the branch is a scalar one.
struct LamReg0 {
  float* p;
  int bd;
  float& l0;
  __device__ __forceinline__ float get(int r) const {
    if (r > 0) return p[(r - 1) * bd];
    return l0;
  }
  __device__ __forceinline__ void set(int r, float v) const {
    if (r > 0) p[(r - 1) * bd] = v;
    else l0 = v;
  }
};

// One predictor-corrector step s over rates r >= s.  `load(r)` gives the
// start log-rate, issued PF rates ahead of its use (PF = 2 for LDS sources; a
// deeper queue hides HBM latency when the state comes from global memory),
// `store(r, l)` receives the end log-rate; e / ep return the final drift
// sums; cs = row s of the step table (fixed drifts of step s).
//
// The rate loop is unrolled by PF and slot u of the queue always serves rates
// with (r - s) % PF == u, so no queue register is ever copied: copying a
// register whose load is still in flight waits for that load, which would
// turn the queue back into one load in flight.  Tail loads are clamped to
// the last rate (in bounds, never used).
template <int F, int PF = 2, class Tab, class Load, class Store>
__device__ __forceinline__ void pc_step(const Model& m, int s, Tab cs,
                                        const float (&z)[F], float (&e)[F], float (&ep)[F],
                                        Load load, Store store) {
  // 32-bit offset: keeps the address scalar
  c_cfloat As = m.A + s * m.N * F;
  const auto kt = cs + (m.N - s) * m.N;
  const int cnt = m.N - s;
#pragma unroll
  for (int f = 0; f < F; ++f) e[f] = ep[f] = 0.0f;
  float q[PF];
#pragma unroll
  for (int i = 0; i < PF; ++i) q[i] = load(s + (i < cnt ? i : cnt - 1));
  for (int b = 0; b < cnt; b += PF) {
#pragma unroll
    for (int u = 0; u < PF; ++u) {
      if (b + u >= cnt) break;
      const int r = s + b + u;
      const float l = q[u];
      float a[F];
#pragma unroll
      for (int f = 0; f < F; ++f) a[f] = As[r * F + f];
      const float kr = kt[r];
      const float g = drift_coeff(l, kr);
      float mu = 0.0f, dw = 0.0f;
#pragma unroll
      for (int f = 0; f < F; ++f) {
        e[f] += g * a[f];
        mu += a[f] * e[f];
        dw += a[f] * z[f];
      }
      const float lh = l + mu + cs[r] + dw;
      const float gh = drift_coeff(lh, kr);
      float muh = 0.0f;
#pragma unroll
      for (int f = 0; f < F; ++f) {
        ep[f] += gh * a[f];
        muh += a[f] * ep[f];
      }
      store(r, lh + 0.5f * (muh - mu));
      // refill after l's last use, so the load can land in l's register
      q[u] = load(s + (b + u + PF < cnt ? b + u + PF : cnt - 1));
    }
  }
}

// Reverse of PC step s over the rates N-1..s.  On entry lam.get(r) holds
// the adjoint of the end-of-step log-rate r, on exit that of the
// start-of-step one (lam.set).  e / ep are the step's final drift sums (un-accumulated
// here), z its variates, cs row s of the step table; `src(r)` gives the start
// log-rate r (issued PF rates ahead).  `fixing(lp, x, f, 1/(1+tau f))` adds
// the product's adjoint of the rate fixing at the end of the step (r == s) to
// lp; `vega(r, w)` receives w = sum_f Abar(s,r,f) A(s,r,f) for the step, which
// contracts to scalars the reverse already has (the mu^, mu, dW and
// fixed-drift terms).  E, E^ are the adjoint running sums.
template <int F, int PF, class Tab, class Src, class Lam, class Fixing, class Vega>
__device__ __forceinline__ void reverse_step(const Model& m, int s, Tab cs, Src src,
                                             const Lam& lam, const float (&z)[F],
                                             float (&e)[F], float (&ep)[F], Fixing fixing,
                                             Vega vega) {
  const int N = m.N;
  // 32-bit offset: keeps the address scalar
  c_cfloat As = m.A + s * N * F;
  const auto kt = cs + (N - s) * N;
  const int cnt = N - s;
  float E[F], Eh[F];
#pragma unroll
  for (int f = 0; f < F; ++f) E[f] = Eh[f] = 0.0f;
  // start log-rates issued PF rates ahead into fixed slots (rate loop
  // unrolled by PF, as in pc_step), adjoints (LDS) one rate ahead
  float q[PF];
#pragma unroll
  for (int i = 0; i < PF; ++i) q[i] = src(N - 1 - (i < cnt ? i : cnt - 1));
  float lp_next = lam.get(N - 1);
  for (int b = 0; b < cnt; b += PF) {
#pragma unroll
    for (int u = 0; u < PF; ++u) {
      if (b + u >= cnt) break;
      const int r = N - 1 - b - u;
      const float l = q[u];
      float lp = lp_next;
      if (r > s) lp_next = lam.get(r - 1);
      float a[F];
#pragma unroll
      for (int f = 0; f < F; ++f) a[f] = As[r * F + f];
      const float kr = kt[r];
      const float g = drift_coeff(l, kr);
      float mu = 0.0f, muh = 0.0f, dw = 0.0f;
#pragma unroll
      for (int f = 0; f < F; ++f) {
        mu += a[f] * e[f];
        muh += a[f] * ep[f];
        dw += a[f] * z[f];
      }
      const float c = cs[r];
      const float lh = l + mu + c + dw;
      const float gh = drift_coeff(lh, kr);
      if (r == s) {
        const float tau = m.taus[s];
        const float xs = fast_exp(lh + 0.5f * (muh - mu));
        const float fs = xs - m.disp[s];
        fixing(lp, xs, fs, fast_rcp(1.0f + tau * fs));
      }
      float ghb = 0.0f;
#pragma unroll
      for (int f = 0; f < F; ++f) {
        Eh[f] += 0.5f * lp * a[f];
        ghb += a[f] * Eh[f];
      }
      const float lhat = lp + ghb * (gh - gh * gh);  // dgh/dlh = gh (1 - gh)
      const float mub = lhat - 0.5f * lp;
      float gb = 0.0f;
#pragma unroll
      for (int f = 0; f < F; ++f) {
        E[f] += mub * a[f];
        gb += a[f] * E[f];
      }
      lam.set(r, lhat + gb * (g - g * g));
      vega(r, 0.5f * lp * muh + gh * ghb + mub * mu + g * gb + lhat * (dw + 2.0f * c));
#pragma unroll
      for (int f = 0; f < F; ++f) {
        e[f] -= g * a[f];
        ep[f] -= gh * a[f];
      }
      q[u] = src(N - 1 - (b + u + PF < cnt ? b + u + PF : cnt - 1));
    }
  }
}

template <int F>
__device__ __forceinline__ void load_z(const Model& m, int s, int p, float (&z)[F]) {
  const float* zs = m.z + static_cast<size_t>(s) * F * m.P + p;
#pragma unroll
  for (int f = 0; f < F; ++f) z[f] = zs[static_cast<size_t>(f) * m.P];
}

// The model arrays are separate __restrict__ kernel arguments (not members of
// a by-value struct) so the compiler may read them through the scalar cache.
#define KOO_MODEL_PARAMS                                                                  \
  const float* __restrict__ z_, const float* __restrict__ f0_,                            \
      const float* __restrict__ taus_, const float* __restrict__ disp_,                   \
      const float* __restrict__ A_, float strike_, int P_, int N_, int k_
#define KOO_MODEL_LOCAL \
  const Model m{z_, cmem(f0_), cmem(taus_), cmem(disp_), cmem(A_), strike_, P_, N_, k_}

// Batched kernels read the pointers of their device descriptor tables as
// global-memory pointers (address space 1), so accesses through them are
// global, not flat, instructions (flat ones also count against the LDS wait
// counter, which would serialise the state prefetch against the LDS
// traffic).  The host fills the same layout with plain pointers.
#if defined(__HIP_DEVICE_COMPILE__)
#define KOO_GLOBAL __attribute__((address_space(1)))
#else
#define KOO_GLOBAL
#endif
typedef const float KOO_GLOBAL g_cfloat;
typedef const double KOO_GLOBAL g_cdouble;
typedef float KOO_GLOBAL g_float;
typedef double KOO_GLOBAL g_double;
typedef int KOO_GLOBAL g_int;
// Address space 1 -> generic (the compiler keeps track that it is global).
template <class T>
__device__ __forceinline__ T* gmem(T KOO_GLOBAL* p) {
  return (T*)p;
}

// Batched launches number their blocks item after item; the item owning
// `block` is the last one whose first block (the member `first`) is <= block
// (binary search, uniform per block).
template <class Item>
__device__ __forceinline__ int find_block_owner(const Item* items, int n, int block,
                                                int Item::*first) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (items[mid].*first <= block) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Rate-count class of a product in a batched portfolio (launches are per
// class so LDS is reserved for the class's largest curve).
inline int rate_class(int rates) { return rates <= 16 ? 0 : (rates <= 24 ? 1 : (rates <= 32 ? 2 : 3)); }

// out[col] = (sum_b part[b*ncols + col]) * scale, compensated; called by a
// whole block (k_finalize_cols runs one block per column).  Agent: the
// partials are read with agent-coherent loads.
template <bool Agent = false>
__device__ __forceinline__ void finalize_col(const double* __restrict__ part, int nb, int ncols,
                                             int col, double scale, double* __restrict__ out) {
  __shared__ double sv[16], sc[16];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  // per lane: Neumaier over a strided slice of the blocks
  double s = 0.0, c = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) {
    const double x = Agent ? load_agent(part + static_cast<size_t>(b) * ncols + col)
                           : part[static_cast<size_t>(b) * ncols + col];
    const double t = s + x;
    c += (fabs(s) >= fabs(x)) ? (s - t) + x : (x - t) + s;
    s = t;
  }
  // wavefront butterfly with error-free additions (TwoSum), then the waves
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const double s2 = __shfl_xor(s, o, 64), c2 = __shfl_xor(c, o, 64);
    double t, e;
    two_sum(s, s2, t, e);
    s = t;
    c += c2 + e;
  }
  if (lane == 0) {
    sv[wave] = s;
    sc[wave] = c;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double S = 0.0, C = 0.0;
    for (int i = 0; i < nw; ++i) {
      double t, e;
      two_sum(S, sv[i], t, e);
      S = t;
      C += sc[i] + e;
    }
    out[col] = (S + C) * scale;
  }
}

__global__ void __launch_bounds__(256)
    k_finalize_cols(const double* __restrict__ part, int nb, int ncols, double scale,
                    double* __restrict__ out) {
  finalize_col(part, nb, ncols, blockIdx.x, scale, out);
}

}  // namespace
}  // namespace lmm
}  // namespace koo
