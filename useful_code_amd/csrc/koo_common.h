// koo_common.h -- shared device helpers and host error plumbing for the
// gfx950 kernels of the LMM pricing library.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/koo_c_abi.h"

namespace koo {

// Records a message for koo_last_error() and returns the status.
int set_error(int status, const char* fmt, ...);

// Launch bookkeeping: after a launch, map hipGetLastError to a status.
int check_launch(const char* what);

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// 16-byte vector of floats for streaming loads and stores: the nontemporal
// builtins (global_load/store ... nt) take clang vector types, not float4.
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ f32x4 load_stream(const f32x4* p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ void store_stream(f32x4 v, f32x4* p) { __builtin_nontemporal_store(v, p); }
__device__ __forceinline__ u32x4 load_stream(const u32x4* p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ void store_stream(u32x4 v, u32x4* p) { __builtin_nontemporal_store(v, p); }
// 16-byte alignment of a device pointer (the condition of the vector paths)
inline bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

// Number of compute units of the current device (cached per device).
int device_cu_count();

// LS workspace (ls_workspace): byte offset of the first pass's partial
// slots, and the most blocks they hold a column for (the running values lead
// the workspace).
size_t ls_workspace_xbuf_offset(int paths);
int ls_workspace_xbuf_columns(int paths);

// 1-D launch helper: blocks for n items at `per_block` items each, capped so a
// grid-stride loop covers the rest.
inline unsigned grid_for(int64_t n, int per_block, int64_t cap = 1 << 20) {
  int64_t b = (n + per_block - 1) / per_block;
  if (b < 1) b = 1;
  if (b > cap) b = cap;
  return static_cast<unsigned>(b);
}

// ---------------------------------------------------------------------------
// Normal quantile of a 32-bit integer uniform (Shaw-Brickman rational
// approximation; W. T. Shaw and N. Brickman, "Differential equations for
// Monte Carlo recycling and a GPU-optimized normal quantile", 2009, with the
// published coefficients that inverseCumulatives.h:146-172 also uses).
//
// The reference maps x -> u = float(x) * 2^-32 and evaluates
// z = -log(1 - |2u - 1|) in FP32, which cancels in both tails.  Here the tail
// mass w = 2 min(u, 1-u) is formed from the integer, t = min(x, 2^32 - x),
// before any rounding, so both tails keep full relative precision; x = 0 is
// read as t = 1/2 (u = 2^-33) where the reference produces a non-finite value.
// Away from the tails the two agree to FP32 rounding.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float invnorm_u32(uint32_t x) {
  const float P1 = 1.2533136835212087879f, P2 = 1.9797154223229267471f,
              P3 = 0.80002295072483916762f, P4 = 0.087403248265958578062f,
              P5 = 0.0020751409553756572917f, P6 = 4.744820732427972462e-6f;
  const float Q2 = 2.0795584360534589311f, Q3 = 1.2499328117341603014f,
              Q4 = 0.23668431621373705623f, Q5 = 0.0120098270559197768f,
              Q6 = 0.00010590620919921025259f;
  const bool upper = x >= 0x80000000u;
  const uint32_t t = upper ? (0u - x) : x;
  const float tf = (t == 0u) ? 0.5f : static_cast<float>(t);
  // w = t * 2^-31 lies in [2^-32, 1]: always a normal number, so the
  // hardware log2 (v_log_f32) needs none of logf's range handling;
  // z = -ln(w) = -ln2 * log2(w).
  const float w = tf * 4.656612873077392578125e-10f;
  const float z = __builtin_amdgcn_logf(w) * -0.693147180559945309417f;
  // Horner with the multiply-adds written out, so every kernel that inlines
  // this rounds alike (the same bits from the staged, fused and persistent
  // generators)
  const float p = fmaf(z, fmaf(z, fmaf(z, fmaf(fmaf(P6, z, P5), z, P4), P3), P2), P1);
  const float q = fmaf(z, fmaf(z, fmaf(z, fmaf(fmaf(Q6, z, Q5), z, Q4), Q3), Q2), 1.0f);
  // z p / q by the hardware reciprocal and one Newton correction of the
  // quotient (q >= 1, so nothing overflows or underflows)
  const float n = z * p;
  const float rq = __builtin_amdgcn_rcpf(q);
  const float r0 = n * rq;
  const float r = fmaf(fmaf(-q, r0, n), rq, r0);
  return upper ? r : -r;
}

// Error-free sum: s + e == a + b exactly (Knuth's TwoSum).
__device__ __forceinline__ void two_sum(double a, double b, double& s, double& e) {
  s = a + b;
  const double bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

// Sobol point for Gray-code index n (x = XOR of v[k] over the set bits of
// g = n ^ (n >> 1)); v in LDS.
__device__ __forceinline__ uint32_t sobol_point(const uint32_t* v, uint32_t n) {
  uint32_t g = n ^ (n >> 1);
  uint32_t x = 0;
  while (g) {
    x ^= v[__builtin_ctz(g)];
    g &= g - 1u;
  }
  return x;
}

}  // namespace koo
