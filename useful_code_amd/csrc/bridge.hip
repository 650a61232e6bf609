// bridge.hip -- stage (a)+(b): Brownian bridge (Jaeckel construction on the
// unit grid) with successive differencing, and the fused
// Sobol -> scramble -> normal -> bridge generator, for gfx950.
//
// One lane owns one (path, factor) pair and every global load/store is a
// contiguous 256-byte row of the path-minor layouts ([D][P] in, (S,F,P)
// out).  For S <= 64 steps the construction order is a compile-time constant
// and the S bridge points of a lane live in VGPRs; longer bridges keep them
// in LDS laid out [step][lane] (bank-conflict free).  The construction
// weights travel in the kernel arguments (scalar loads), so a call allocates
// nothing.
#include <array>
#include <cmath>
#include <mutex>
#include <utility>
#include <unordered_map>
#include <vector>

#include "bridge_device.h"

namespace koo {

namespace {

__global__ void __launch_bounds__(kBridgeThreads)
    k_bridge(const BridgeTables T, int steps, int factors, int paths,
             const int* __restrict__ alloc, const float* __restrict__ in,
             float* __restrict__ out) {
  extern __shared__ float lds[];
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= paths) return;
  const int* idx = alloc + blockIdx.y * steps;
  bridge_lane(T, steps, factors, paths, p, blockIdx.y, lds,
              [&](int i) { return in[static_cast<size_t>(idx[i]) * paths + p]; }, out);
}

// Fused generator, general steps: the bridge variate i of factor f is the
// normal quantile of Sobol dimension d = alloc(f, i) at Gray index p + off1,
// scrambled when d < number_scrambled.  Direction numbers of the block's S
// dimensions sit in LDS behind the path storage.
__global__ void __launch_bounds__(kBridgeThreads)
    k_sobol_bridge(const BridgeTables T, int steps, int factors, int paths, uint32_t off1,
                   const uint32_t* __restrict__ dirs, const uint32_t* __restrict__ scr,
                   int number_scrambled, const int* __restrict__ alloc, float* __restrict__ out) {
  extern __shared__ float lds[];
  uint32_t* v = reinterpret_cast<uint32_t*>(lds + steps * blockDim.x);
  const int* idx = alloc + blockIdx.y * steps;
  for (int w = threadIdx.x; w < steps * 32; w += blockDim.x)
    v[w] = dirs[static_cast<size_t>(idx[w >> 5]) * 32 + (w & 31)];
  __syncthreads();
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= paths) return;
  const uint32_t n = static_cast<uint32_t>(p) + off1;
  bridge_lane(T, steps, factors, paths, p, blockIdx.y, lds,
              [&](int i) {
                const int d = idx[i];
                const uint32_t s = d < number_scrambled ? scr[d] : 0u;
                return invnorm_u32(sobol_point(v + 32 * i, n) ^ s);
              },
              out);
}

// ---------------------------------------------------------------------------
// Register-resident bridge for S <= kRegSteps.  The construction order
// depends on S only, so for a compile-time S every index is a constant and
// the S bridge points of a lane stay in VGPRs (no LDS path storage, which
// held occupancy to one or two waves per SIMD).  The stand-alone bridge
// reads its weights from the kernel arguments (uniform, scalar loads); the
// fused generator, which loops over chunks, from an LDS copy (LdsWeights).
// ---------------------------------------------------------------------------
constexpr int kRegSteps = 64;

// The index walk of make_tables (brownian_bridge.cu:639-700) as a constant
// expression.
template <int S>
struct BridgeOrder {
  int16_t bridge[S] = {}, left[S] = {}, right[S] = {};
  constexpr BridgeOrder() {
    int map[S] = {};
    map[S - 1] = 1;
    bridge[0] = S - 1;
    for (int j = 0, i = 1; i < S; ++i) {
      while (map[j]) ++j;
      int k = j;
      while (!map[k]) ++k;
      const int l = j + ((k - 1 - j) >> 1);
      map[l] = i;
      bridge[i] = static_cast<int16_t>(l);
      left[i] = static_cast<int16_t>(j);
      right[i] = static_cast<int16_t>(k);
      j = k + 1;
      if (j >= S) j = 0;
    }
  }
};

// The construction weights of step i, {rw, lw, sd}, from a copy of the
// tables in LDS ({rw, lw, sd, 0} per step) read at an index offset the
// compiler cannot see through.  A loop over chunks that reads the
// kernel-argument tables gets all 3 S weights hoisted into scalar registers,
// more than there are (spilled through vector lanes); the LDS copy is read
// where each weight is used.
struct LdsWeights {
  const f32x4* w;  // LDS
  int z;           // zero, opaque to the compiler
  __device__ __forceinline__ void get(int i, float& rw, float& lw, float& sd) const {
    const f32x4 q = w[i + z];
    rw = q.x;
    lw = q.y;
    sd = q.z;
  }
};

// V is float (one path) or f32x4 (four paths; the scalar weights broadcast).
template <int S, int I, class V, class Z, class W>
__device__ __forceinline__ void bridge_point(V (&w)[S], const W& wt, Z& z) {
  constexpr BridgeOrder<S> M{};
  float rw, lw, sd;
  wt.get(I, rw, lw, sd);
  if constexpr (I == 0) {
    w[S - 1] = sd * z(0);
  } else {
    constexpr int j = M.left[I], k = M.right[I], l = M.bridge[I];
    V v = rw * w[k];
    if constexpr (j != 0) v = bridge_fma(lw, w[j - 1], v);
    w[l] = bridge_fma(sd, z(I), v);
  }
}

template <int S, class V, class Z, class W, int... I>
__device__ __forceinline__ void bridge_points(V (&w)[S], const W& wt, Z& z,
                                              std::integer_sequence<int, I...>) {
  (bridge_point<S, I>(w, wt, z), ...);
}

// In-place form for variates that come from memory: w[bridge(i)] holds
// variate i on entry (so all S loads can be in flight at once, into the
// registers the points end in), and point bridge(i) replaces it; the same
// operations in the same order as bridge_point.
template <int S, int I, class V>
__device__ __forceinline__ void bridge_point_inplace(V (&w)[S], const BridgeTables& T) {
  constexpr BridgeOrder<S> M{};
  if constexpr (I == 0) {
    w[S - 1] = T.sd[0] * w[S - 1];
  } else {
    constexpr int j = M.left[I], k = M.right[I], l = M.bridge[I];
    V v = T.rw[I] * w[k];
    if constexpr (j != 0) v = bridge_fma(T.lw[I], w[j - 1], v);
    w[l] = bridge_fma(T.sd[I], w[l], v);
  }
}

template <int S, class V, class Load, int... I>
__device__ __forceinline__ void bridge_points_loaded(V (&w)[S], const BridgeTables& T,
                                                     Load load,
                                                     std::integer_sequence<int, I...>) {
  constexpr BridgeOrder<S> M{};
  ((w[M.bridge[I]] = load(I)), ...);
  (bridge_point_inplace<S, I>(w, T), ...);
}

// Builds the lane's path from z(i) (bridge variate i, i a constant) and
// stores its increments to (S,F,P) at path p when `live`.
template <int S, class Z, class W>
__device__ __forceinline__ void bridge_store(const W& wt, Z z, int factors, int f, int paths,
                                             int p, bool live, float* __restrict__ out) {
  float w[S];
  bridge_points<S>(w, wt, z, std::make_integer_sequence<int, S>{});
  // the step index offset by the weights' opaque zero, so the S step
  // offsets are formed where they are stored rather than hoisted into
  // scalar registers across the chunk loop
  float prev = 0.0f;
#pragma unroll
  for (int i = 0; i < S; ++i) {
    if (live) out[(static_cast<size_t>(i + wt.z) * factors + f) * paths + p] = w[i] - prev;
    prev = w[i];
  }
}

template <int S>
__global__ void __launch_bounds__(kBridgeThreads)
    k_bridge_reg(const BridgeTables T, int factors, int paths, const int* __restrict__ alloc,
                 const float* __restrict__ in, float* __restrict__ out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= paths) return;
  const int* idx = alloc + blockIdx.y * S;
  const size_t P = static_cast<size_t>(paths);
  float w[S];
  bridge_points_loaded<S>(
      w, T, [&](int i) { return in[static_cast<size_t>(idx[i]) * P + p]; },
      std::make_integer_sequence<int, S>{});
  float prev = 0.0f;
#pragma unroll
  for (int i = 0; i < S; ++i) {
    out[(static_cast<size_t>(i) * factors + blockIdx.y) * P + p] = w[i] - prev;
    prev = w[i];
  }
}

// Four paths per lane (16-byte streaming loads and stores, four times the
// contiguous bytes per wavefront and row) for S <= kRegSteps4.
constexpr int kRegSteps4 = 48;

template <int S>
__global__ void __launch_bounds__(kBridgeThreads)
    k_bridge_reg4(const BridgeTables T, int factors, int paths4, const int* __restrict__ alloc,
                  const f32x4* __restrict__ in, f32x4* __restrict__ out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= paths4) return;
  const size_t P4 = static_cast<size_t>(paths4);
  const int* idx = alloc + blockIdx.y * S;
  f32x4 w[S];
  bridge_points_loaded<S>(
      w, T, [&](int i) { return load_stream(in + static_cast<size_t>(idx[i]) * P4 + p); },
      std::make_integer_sequence<int, S>{});
  f32x4 prev = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int i = 0; i < S; ++i) {
    store_stream(w[i] - prev, out + (static_cast<size_t>(i) * factors + blockIdx.y) * P4 + p);
    prev = w[i];
  }
}

// Fused generator on the register bridge.  Each wavefront walks a contiguous
// run of 64-path chunks of one factor.  With n = G + l (G a multiple of 64,
// l < 64) the Gray code splits, gray(n) = gray(G) ^ gray(l), so the
// scrambled Sobol integer is X(G) ^ T(l) with X(G) = scrambler ^ XOR of the
// direction numbers over the set bits of gray(G).  A lane's l is the same in
// every chunk; T(l) sits in LDS.  X(G) lives one dimension per lane: from
// G = 64m to 64(m+1) gray flips bits 5 and 6 + ctz(m+1), so advancing costs
// two XORs; a chunk (lanes at n = 64c + off1 + lane) spans G and G + 64, so
// both are formed and read out by readlane.
template <int S>
__global__ void __launch_bounds__(kBridgeThreads)
    k_sobol_bridge_reg(const BridgeTables T, int factors, int paths, uint32_t off1,
                       const uint32_t* __restrict__ dirs, const uint32_t* __restrict__ scr,
                       int number_scrambled, const int* __restrict__ alloc,
                       float* __restrict__ out) {
  __shared__ uint32_t vall[S * 32];  // direction numbers [bridge variate][bit]
  __shared__ uint32_t tlo[S * 64];   // T(l) [bridge variate][l]
  __shared__ f32x4 wts[S];           // construction weights {rw, lw, sd, 0}
  const int f = blockIdx.y;
  const int* idx = alloc + f * S;
  for (int q = threadIdx.x; q < S * 32; q += blockDim.x)
    vall[q] = dirs[static_cast<size_t>(idx[q >> 5]) * 32 + (q & 31)];
  for (int q = threadIdx.x; q < S; q += blockDim.x) wts[q] = f32x4{T.rw[q], T.lw[q], T.sd[q], 0.0f};
  __syncthreads();
  for (int q = threadIdx.x; q < S * 64; q += blockDim.x) {
    const uint32_t l = static_cast<uint32_t>(q & 63), g = l ^ (l >> 1);
    const uint32_t* vv = vall + (q >> 6) * 32;
    uint32_t x = 0;
#pragma unroll
    for (int bit = 0; bit < 6; ++bit)
      if ((g >> bit) & 1u) x ^= vv[bit];
    tlo[q] = x;
  }
  __syncthreads();

  const int lane = threadIdx.x & 63;
  const int64_t waves = static_cast<int64_t>(gridDim.x) * (blockDim.x >> 6);
  const int64_t wv = static_cast<int64_t>(blockIdx.x) * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int64_t chunks = (static_cast<int64_t>(paths) + 63) >> 6;
  const int c0 = static_cast<int>(wv * chunks / waves);
  const int c1 = static_cast<int>((wv + 1) * chunks / waves);
  if (c0 >= c1) return;

  // this lane's dimension for the X(G) bookkeeping
  const int dl = lane < S ? lane : S - 1;
  const uint32_t* vd = vall + dl * 32;
  const uint32_t r = off1 & 63u;
  const bool upper = r + static_cast<uint32_t>(lane) >= 64u;  // lane sits in G + 64
  const uint32_t l = (r + static_cast<uint32_t>(lane)) & 63u;
  uint32_t m = static_cast<uint32_t>(c0) + (off1 >> 6);  // G = 64 m for chunk c0
  uint32_t xc;
  {
    const int d = idx[dl];
    xc = d < number_scrambled ? scr[d] : 0u;
    const uint32_t G = m << 6, g = G ^ (G >> 1);
    for (int b = 5; b < 32; ++b)
      if ((g >> b) & 1u) xc ^= vd[b];
  }
  const uint32_t v5 = vd[5];
  for (int c = c0; c < c1; ++c, ++m) {
    const int t = 6 + __builtin_ctz(m + 1u);
    const uint32_t xn = xc ^ v5 ^ (t < 32 ? vd[t] : 0u);
    const int p = c * 64 + lane;
    auto z = [&](int i) {
      const uint32_t xa = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(xc), i));
      const uint32_t xb = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(xn), i));
      return invnorm_u32((upper ? xb : xa) ^ tlo[i * 64 + l]);
    };
    // every lane computes (the readlanes see whole-wave values); only the
    // lanes inside the path range store.  The weights come from LDS at an
    // offset made opaque each chunk (a zero from s_mov), so their reads stay
    // in the loop.
    int zero;
    asm volatile("s_mov_b32 %0, 0" : "=s"(zero));
    bridge_store<S>(LdsWeights{wts, zero}, z, factors, f, paths, p, p < paths, out);
    xc = xn;
  }
}

// Launch tables for every S in [1, kRegSteps].
using RegBridgeFn = void (*)(const BridgeTables, int, int, const int*, const float*, float*);
using RegGenFn = void (*)(const BridgeTables, int, int, uint32_t, const uint32_t*,
                          const uint32_t*, int, const int*, float*);

template <int... I>
constexpr auto reg_bridge_table(std::integer_sequence<int, I...>) {
  return std::array<RegBridgeFn, sizeof...(I)>{&k_bridge_reg<I + 1>...};
}
template <int... I>
constexpr auto reg_gen_table(std::integer_sequence<int, I...>) {
  return std::array<RegGenFn, sizeof...(I)>{&k_sobol_bridge_reg<I + 1>...};
}
using RegBridge4Fn = void (*)(const BridgeTables, int, int, const int*, const f32x4*, f32x4*);
template <int... I>
constexpr auto reg_bridge4_table(std::integer_sequence<int, I...>) {
  return std::array<RegBridge4Fn, sizeof...(I)>{&k_bridge_reg4<I + 1>...};
}
const auto kRegBridge = reg_bridge_table(std::make_integer_sequence<int, kRegSteps>{});
const auto kRegBridge4 = reg_bridge4_table(std::make_integer_sequence<int, kRegSteps4>{});
const auto kRegGen = reg_gen_table(std::make_integer_sequence<int, kRegSteps>{});

// ---------------------------------------------------------------------------
// Power-of-two bridge (the reference's legacy BrownianBridge<D>,
// gold/Bridge_gold.h:168-252, GPU Bridge_gpu.cu:58-195): S = 2^n points;
// the last point first (scale 2^(n/2)), then the one-sided points
// 2^(n-i-1) - 1 from their right neighbour, then level by level the
// midpoints between two built points; point k uses variate var(k), the
// order in which the construction visits it.  Input and output are
// [S][vectors] (vector-minor), the output differenced.
// ---------------------------------------------------------------------------
constexpr int kPow2MaxLog2 = 8;

struct Pow2Tables {
  float first;                   // 2^(n/2)
  float one[kPow2MaxLog2];       // one-sided step i: sqrt(2^(n-i-1) / 2)
  float mid[kPow2MaxLog2 + 1];   // two-sided level i: sqrt(2^(n-i) / 2)
};

// Point -> variate of the construction (gold/Bridge_gold.h:237-251).
template <int LOG2>
struct Pow2Order {
  int16_t var[1 << LOG2] = {};
  constexpr Pow2Order() {
    int c = 0;
    for (int i = 0; i <= LOG2; ++i) {
      const int m = 1 << (LOG2 - i);
      for (int j = 1; j <= (1 << i); j += 2) var[j * m - 1] = static_cast<int16_t>(c++);
    }
  }
};

__device__ __forceinline__ int pow2_var(int LOG2, int k) {
  // position of point k in the construction: k + 1 = j 2^(n-i) with j odd
  const int t = __builtin_ctz(static_cast<unsigned>(k + 1));  // n - i
  const int i = LOG2 - t;
  const int j = (k + 1) >> t;
  return (i == 0) ? 0 : (1 << (i - 1)) + (j >> 1);
}

template <int LOG2>
__global__ void __launch_bounds__(kBridgeThreads)
    k_bridge_pow2_reg(const Pow2Tables T, int vectors, const float* __restrict__ in,
                      float* __restrict__ out) {
  constexpr int S = 1 << LOG2;
  constexpr Pow2Order<LOG2> M{};
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= vectors) return;
  const size_t n = static_cast<size_t>(vectors);
  auto z = [&](int k) { return in[static_cast<size_t>(M.var[k]) * n + p]; };
  float w[S];
  w[S - 1] = z(S - 1) * T.first;
#pragma unroll
  for (int i = 0; i < LOG2; ++i) {
    const int k = (1 << (LOG2 - i - 1)) - 1, right = (1 << (LOG2 - i)) - 1;
    w[k] = 0.5f * w[right] + T.one[i] * z(k);
  }
#pragma unroll
  for (int i = 2; i <= LOG2; ++i) {
    const int m = 1 << (LOG2 - i);
#pragma unroll
    for (int j = 3; j < (1 << i); j += 2) {
      const int k = j * m - 1;
      w[k] = 0.5f * w[k + m] + 0.5f * w[k - m] + T.mid[i] * z(k);
    }
  }
  out[p] = w[0];
#pragma unroll
  for (int k = 1; k < S; ++k) out[static_cast<size_t>(k) * n + p] = w[k] - w[k - 1];
}

// Same construction with the points in LDS [point][lane] (S = 128, 256).
__global__ void __launch_bounds__(64)
    k_bridge_pow2_lds(const Pow2Tables T, int log2, int vectors, const float* __restrict__ in,
                      float* __restrict__ out) {
  extern __shared__ float lds[];
  const int S = 1 << log2;
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= vectors) return;
  const int bd = blockDim.x;
  float* w = lds + threadIdx.x;  // w[k * bd]
  const size_t n = static_cast<size_t>(vectors);
  auto z = [&](int k) { return in[static_cast<size_t>(pow2_var(log2, k)) * n + p]; };
  w[(S - 1) * bd] = z(S - 1) * T.first;
  for (int i = 0; i < log2; ++i) {
    const int k = (1 << (log2 - i - 1)) - 1, right = (1 << (log2 - i)) - 1;
    w[k * bd] = 0.5f * w[right * bd] + T.one[i] * z(k);
  }
  for (int i = 2; i <= log2; ++i) {
    const int m = 1 << (log2 - i);
    for (int j = 3; j < (1 << i); j += 2) {
      const int k = j * m - 1;
      w[k * bd] = 0.5f * w[(k + m) * bd] + 0.5f * w[(k - m) * bd] + T.mid[i] * z(k);
    }
  }
  float prev = 0.0f;
  for (int k = 0; k < S; ++k) {
    const float v = w[k * bd];
    out[static_cast<size_t>(k) * n + p] = v - prev;
    prev = v;
  }
}

// Multi-factor reordering (Bridge_multi_gpu.cu:67-112 with the orderings of
// BrownianBridgeMultiDim, gold/Bridge_gold.h:291-331): Sobol dimension i of
// path p goes to (bridge variate v(i), factor f(i)) of the (S, P, F) array
// the power-of-two bridge then transforms as S x (P F) vectors.  The
// ordering is walked in scalar registers (the same for every lane), so no
// table is needed.  ordering: 0 row, 1 column, 2 triangular.
__global__ void __launch_bounds__(256)
    k_bridge_reorder(const float* __restrict__ in, float* __restrict__ out, int points,
                     int factors, int paths, int ordering) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= paths) return;
  const size_t P = static_cast<size_t>(paths);
  const int N = points * factors;
  int crow = 0, ccol = 0;
  for (int i = 0; i < N; ++i) {
    int v, f;
    if (ordering == 1) {
      v = i % points;
      f = i / points;
    } else if (ordering == 0) {
      v = i / factors;
      f = i % factors;
    } else {
      v = crow;
      f = ccol;
      if (ccol > 0 && crow < points - 1) {
        --ccol;
        ++crow;
      } else {
        const int rpc = crow + ccol + 1;
        ccol = rpc < factors - 1 ? rpc : factors - 1;
        crow = rpc - ccol;
      }
    }
    out[(static_cast<size_t>(v) * P + p) * factors + f] = in[static_cast<size_t>(i) * P + p];
  }
}

using Pow2Fn = void (*)(const Pow2Tables, int, const float*, float*);
template <int... I>
constexpr auto pow2_table(std::integer_sequence<int, I...>) {
  return std::array<Pow2Fn, sizeof...(I)>{&k_bridge_pow2_reg<I>...};
}
constexpr int kPow2RegMaxLog2 = 6;
const auto kPow2Reg = pow2_table(std::make_integer_sequence<int, kPow2RegMaxLog2 + 1>{});

// The reference's float tables (gold/Bridge_gold.h:168-200): scalars are
// formed in double and rounded to float.
Pow2Tables make_pow2_tables(int log2) {
  Pow2Tables T{};
  T.first = static_cast<float>(std::pow(2.0, 0.5 * log2));
  for (int i = 0; i < log2; ++i) {
    const float t = static_cast<float>(std::pow(2.0, log2 - i - 1.0));
    T.one[i] = static_cast<float>(std::sqrt(0.5 * t));
  }
  for (int i = 2; i <= log2; ++i) {
    const int m = 1 << (log2 - i);
    T.mid[i] = static_cast<float>(std::sqrt(m * 0.5));
  }
  return T;
}

}  // namespace
}  // namespace koo

using namespace koo;

// Diagonal allocation; the walk of gold/Bridge_allocator_gold.cpp:38-76.
extern "C" int bridge_allocate_diagonal(int factors, int steps, int* indices) {
  if (factors < 1 || steps < 1 || !indices)
    return set_error(KOO_ERR_INVALID, "bridge_allocate_diagonal: bad arguments");
  int i0 = 0, j0 = 0, i = 0, j = 0, counter = 0;
  while (counter < factors * steps) {
    indices[i * steps + j] = counter++;
    if (i == 0 || j == steps - 1) {
      if (i0 < factors - 1) {
        i0 += 1;
        j0 = 0;
      } else {
        i0 = factors - 1;
        j0 += 1;
      }
      i = i0;
      j = j0;
    } else {
      i -= 1;
      j += 1;
    }
  }
  return KOO_OK;
}

// Lanes per block so that the LDS path storage stays within 64 KiB.
static int bridge_threads(int steps) { return steps <= 48 ? 256 : (steps <= 96 ? 128 : 64); }

static int bridge_check(const char* who, int steps, int paths, int factors) {
  if (steps < 1 || steps > kBridgeMaxSteps)
    return set_error(KOO_ERR_INVALID, "%s: steps %d outside [1,%d]", who, steps, kBridgeMaxSteps);
  if (paths < 0 || factors < 1 || factors > 65535)
    return set_error(KOO_ERR_INVALID, "%s: bad paths/factors (%d, %d)", who, paths, factors);
  return KOO_OK;
}

extern "C" int brownian_bridge_gpu(int steps, const int* allocation_indices, const float* in,
                                   float* out, int paths, int factors, void* stream) {
  if (int e = bridge_check("brownian_bridge_gpu", steps, paths, factors)) return e;
  if (paths == 0) return KOO_OK;
  if (!allocation_indices || !in || !out)
    return set_error(KOO_ERR_INVALID, "brownian_bridge_gpu: null pointer");
  if (steps <= kRegSteps4 && paths % 4 == 0 && aligned16(in) && aligned16(out)) {
    const int paths4 = paths / 4;
    hipLaunchKernelGGL(kRegBridge4[steps - 1],
                       dim3((paths4 + kBridgeThreads - 1) / kBridgeThreads, factors),
                       dim3(kBridgeThreads), 0, as_stream(stream), tables_for(steps), factors,
                       paths4, allocation_indices, reinterpret_cast<const f32x4*>(in),
                       reinterpret_cast<f32x4*>(out));
    return check_launch("brownian_bridge_gpu");
  }
  if (steps <= kRegSteps) {
    hipLaunchKernelGGL(kRegBridge[steps - 1], dim3((paths + kBridgeThreads - 1) / kBridgeThreads,
                                                   factors),
                       dim3(kBridgeThreads), 0, as_stream(stream), tables_for(steps), factors, paths,
                       allocation_indices, in, out);
    return check_launch("brownian_bridge_gpu");
  }
  const int threads = bridge_threads(steps);
  const size_t lds = sizeof(float) * steps * threads;
  hipLaunchKernelGGL(k_bridge, dim3((paths + threads - 1) / threads, factors), dim3(threads), lds,
                     as_stream(stream), tables_for(steps), steps, factors, paths,
                     allocation_indices, in, out);
  return check_launch("brownian_bridge_gpu");
}

extern "C" int sobol_bridge_gpu(int paths, unsigned path_offset, int steps, int factors,
                                const unsigned int* d_directions, const unsigned int* scramblers,
                                int number_scrambled, const int* allocation_indices, float* out,
                                void* stream) {
  if (int e = bridge_check("sobol_bridge_gpu", steps, paths, factors)) return e;
  if (paths == 0) return KOO_OK;
  if (!d_directions || !allocation_indices || !out || (number_scrambled > 0 && !scramblers))
    return set_error(KOO_ERR_INVALID, "sobol_bridge_gpu: null pointer");
  if (static_cast<uint64_t>(path_offset) + 1u + static_cast<uint64_t>(paths) > (1ull << 32))
    return set_error(KOO_ERR_INVALID, "sobol_bridge_gpu: path_offset + paths exceeds 2^32 - 1");
  if (steps <= kRegSteps) {
    // One block set that is resident at once (blocks per CU from the
    // occupancy calculator), split evenly over the factors; each wavefront
    // then walks its own run of chunks, so there is no tail of late blocks.
    static std::array<int, kRegSteps> resident{};
    static std::mutex mu;
    int per_cu;
    {
      std::lock_guard<std::mutex> lock(mu);
      if (!resident[steps - 1]) {
        int nb = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &nb, reinterpret_cast<const void*>(kRegGen[steps - 1]), kBridgeThreads, 0) !=
                hipSuccess ||
            nb < 1)
          nb = 1;
        resident[steps - 1] = nb;
      }
      per_cu = resident[steps - 1];
    }
    const int64_t chunks = (static_cast<int64_t>(paths) + 63) / 64;
    int64_t per_factor = (static_cast<int64_t>(per_cu) * device_cu_count()) / factors;
    const int64_t most = (chunks + kBridgeThreads / 64 - 1) / (kBridgeThreads / 64);
    if (per_factor > most) per_factor = most;
    if (per_factor < 1) per_factor = 1;
    hipLaunchKernelGGL(kRegGen[steps - 1], dim3(static_cast<unsigned>(per_factor), factors),
                       dim3(kBridgeThreads), 0, as_stream(stream), tables_for(steps), factors,
                       paths, path_offset + 1u, d_directions, scramblers, number_scrambled,
                       allocation_indices, out);
    return check_launch("sobol_bridge_gpu");
  }
  const int threads = bridge_threads(steps);
  const size_t lds = sizeof(float) * steps * threads + sizeof(uint32_t) * 32 * steps;
  hipLaunchKernelGGL(k_sobol_bridge, dim3((paths + threads - 1) / threads, factors), dim3(threads),
                     lds, as_stream(stream), tables_for(steps), steps,
                     factors, paths, path_offset + 1u, d_directions, scramblers,
                     number_scrambled, allocation_indices, out);
  return check_launch("sobol_bridge_gpu");
}

// Bridge_gpu.h:11 (brownianBridgeGPU_constant_memory): power-of-two bridge of
// n_vectors vectors of 2^n_poweroftwo points, [points][vectors] in and out.
extern "C" int brownianBridgeGPU_constant_memory(int n_vectors, int n_poweroftwo,
                                                 const float* input, float* output,
                                                 void* stream) {
  if (n_vectors < 0 || n_poweroftwo < 0 || n_poweroftwo > kPow2MaxLog2)
    return set_error(KOO_ERR_INVALID,
                     "brownianBridgeGPU_constant_memory: n_vectors %d, n_poweroftwo %d outside "
                     "[0,%d]",
                     n_vectors, n_poweroftwo, kPow2MaxLog2);
  if (n_vectors == 0) return KOO_OK;
  if (!input || !output)
    return set_error(KOO_ERR_INVALID, "brownianBridgeGPU_constant_memory: null pointer");
  const Pow2Tables T = make_pow2_tables(n_poweroftwo);
  hipStream_t st = as_stream(stream);
  if (n_poweroftwo <= kPow2RegMaxLog2) {
    hipLaunchKernelGGL(kPow2Reg[n_poweroftwo],
                       dim3((n_vectors + kBridgeThreads - 1) / kBridgeThreads),
                       dim3(kBridgeThreads), 0, st, T, n_vectors, input, output);
  } else {
    const int threads = 64;
    const size_t lds = sizeof(float) * (static_cast<size_t>(1) << n_poweroftwo) * threads;
    hipLaunchKernelGGL(k_bridge_pow2_lds, dim3((n_vectors + threads - 1) / threads),
                       dim3(threads), lds, st, T, n_poweroftwo, n_vectors, input, output);
  }
  return check_launch("brownianBridgeGPU_constant_memory");
}

// Bridge_multi_gpu.h:12-20 (brownianBridgeMultiGPUReorder): [D][P] Sobol
// dimensions to the (S, P, F) arrangement of bridge variates and factors,
// D = 2^n_poweroftwo * factors.  allocator: 0 row, 1 column, 2 triangular
// (BrownianBridgeMultiDim<float>::ordering).  useTextures is ignored.
extern "C" int brownianBridgeMultiGPUReorder(const float* input, float* output, int n_poweroftwo,
                                             int factors, int number_paths, int allocator,
                                             int /*useTextures*/, void* stream) {
  if (n_poweroftwo < 0 || n_poweroftwo > kPow2MaxLog2 || factors < 1 || number_paths < 0 ||
      allocator < 0 || allocator > 2)
    return set_error(KOO_ERR_INVALID,
                     "brownianBridgeMultiGPUReorder: bad arguments (n %d, factors %d, paths %d, "
                     "allocator %d)",
                     n_poweroftwo, factors, number_paths, allocator);
  if (number_paths == 0) return KOO_OK;
  if (!input || !output)
    return set_error(KOO_ERR_INVALID, "brownianBridgeMultiGPUReorder: null pointer");
  hipLaunchKernelGGL(k_bridge_reorder, dim3((number_paths + 255) / 256), dim3(256), 0,
                     as_stream(stream), input, output, 1 << n_poweroftwo, factors, number_paths,
                     allocator);
  return check_launch("brownianBridgeMultiGPUReorder");
}
