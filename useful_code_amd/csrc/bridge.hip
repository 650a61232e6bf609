// bridge.hip -- stage (a)+(b): Brownian bridge (Jaeckel construction on the
// unit grid) with successive differencing, and the fused
// Sobol -> scramble -> normal -> bridge generator, for gfx950.
//
// One lane owns one (path, factor) pair; the S bridge points of that pair
// live in LDS laid out [step][lane] so every LDS access of a wavefront is
// bank-conflict free, and every global load/store is a contiguous 256-byte
// row of the path-minor layouts ([D][P] in, (S,F,P) out).  The construction
// tables travel in the kernel arguments (scalar loads), so a call allocates
// nothing.
#include <mutex>
#include <unordered_map>
#include <vector>

#include "koo_common.h"

namespace koo {

namespace {

constexpr int kBridgeMaxSteps = 128;

// Construction schedule of brownian_bridge::initialize (brownian_bridge.cu:
// 639-700) on t_i = i + 1 (brownian_bridge.cu:359-362), kept in float with
// the same operations so the weights are the reference's bit for bit.
struct BridgeTables {
  int16_t bridge[kBridgeMaxSteps];
  int16_t left[kBridgeMaxSteps];
  int16_t right[kBridgeMaxSteps];
  float lw[kBridgeMaxSteps];
  float rw[kBridgeMaxSteps];
  float sd[kBridgeMaxSteps];
};

BridgeTables make_tables(int size) {
  BridgeTables T{};
  std::vector<float> t(size);
  for (int i = 0; i < size; ++i) t[i] = static_cast<float>(i + 1);
  std::vector<int> map(size, 0);
  map[size - 1] = 1;
  T.bridge[0] = static_cast<int16_t>(size - 1);
  T.sd[0] = std::sqrt(t[size - 1]);
  T.lw[0] = T.rw[0] = 0.0f;
  for (int j = 0, i = 1; i < size; ++i) {
    while (map[j]) ++j;
    int k = j;
    while (!map[k]) ++k;
    const int l = j + ((k - 1 - j) >> 1);
    map[l] = i;
    T.bridge[i] = static_cast<int16_t>(l);
    T.left[i] = static_cast<int16_t>(j);
    T.right[i] = static_cast<int16_t>(k);
    if (j != 0) {
      T.lw[i] = (t[k] - t[l]) / (t[k] - t[j - 1]);
      T.rw[i] = (t[l] - t[j - 1]) / (t[k] - t[j - 1]);
      T.sd[i] = std::sqrt(((t[l] - t[j - 1]) * (t[k] - t[l])) / (t[k] - t[j - 1]));
    } else {
      T.lw[i] = (t[k] - t[l]) / t[k];
      T.rw[i] = t[l] / t[k];
      T.sd[i] = std::sqrt(t[l] * (t[k] - t[l]) / t[k]);
    }
    j = k + 1;
    if (j >= size) j = 0;
  }
  return T;
}

const BridgeTables& tables_for(int steps) {
  static std::mutex mu;
  static std::unordered_map<int, BridgeTables> cache;
  std::lock_guard<std::mutex> lock(mu);
  auto it = cache.find(steps);
  if (it == cache.end()) it = cache.emplace(steps, make_tables(steps)).first;
  return it->second;
}

constexpr int kBridgeThreads = 256;

// Builds one factor's path for this lane's path index from the variate
// source `z(i)` (bridge variate i), differences it and stores (S,F,P).
template <class Source>
__device__ __forceinline__ void bridge_lane(const BridgeTables& T, int steps, int factors,
                                            int paths, int p, float* lds, Source z,
                                            float* __restrict__ out) {
  const int f = blockIdx.y;
  const int lane = threadIdx.x;
  float* path = lds + lane;  // path[i * blockDim.x]
  const int bd = blockDim.x;
  path[(steps - 1) * bd] = T.sd[0] * z(0);
  for (int i = 1; i < steps; ++i) {
    const int j = T.left[i], k = T.right[i], l = T.bridge[i];
    float v = T.rw[i] * path[k * bd];
    if (j != 0) v = T.lw[i] * path[(j - 1) * bd] + v;
    path[l * bd] = v + T.sd[i] * z(i);
  }
  float prev = 0.0f;
  for (int i = 0; i < steps; ++i) {
    const float w = path[i * bd];
    out[(static_cast<size_t>(i) * factors + f) * paths + p] = w - prev;
    prev = w;
  }
}

__global__ void __launch_bounds__(kBridgeThreads)
    k_bridge(const BridgeTables T, int steps, int factors, int paths,
             const int* __restrict__ alloc, const float* __restrict__ in,
             float* __restrict__ out) {
  extern __shared__ float lds[];
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= paths) return;
  const int* idx = alloc + blockIdx.y * steps;
  bridge_lane(T, steps, factors, paths, p, lds,
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
  bridge_lane(T, steps, factors, paths, p, lds,
              [&](int i) {
                const int d = idx[i];
                const uint32_t s = d < number_scrambled ? scr[d] : 0u;
                return invnorm_u32(sobol_point(v + 32 * i, n) ^ s);
              },
              out);
}

// Fused generator, steps <= kTableSteps: no per-point bit loop.  With
// n = G + l, G a multiple of 64 and l < 64, the Gray code splits,
// gray(n) = gray(G) ^ gray(l), so the Sobol integer is x(n) = X(G) ^ T(l).
// A chunk of blockDim consecutive paths spans at most blockDim/64 + 1
// multiples of 64.  Each block keeps the direction numbers of its S
// dimensions and the 64-entry tables T(l) in LDS for its whole life and walks
// chunks of paths; per chunk it builds X(G) for those groups (scrambler
// folded in; each one a 32-lane XOR reduction over direction numbers read
// from LDS).  A point then costs two LDS reads and one XOR.
constexpr int kTableSteps = 64;
constexpr int kMaxGroups = 8;

__global__ void __launch_bounds__(kBridgeThreads)
    k_sobol_bridge_tab(const BridgeTables T, int steps, int factors, int paths, uint32_t off1,
                       const uint32_t* __restrict__ dirs, const uint32_t* __restrict__ scr,
                       int number_scrambled, const int* __restrict__ alloc,
                       float* __restrict__ out) {
  extern __shared__ float lds[];
  const int bd = blockDim.x;
  uint32_t* vall = reinterpret_cast<uint32_t*>(lds + steps * bd);  // [steps][32]
  uint32_t* tlo = vall + steps * 32;                                 // [steps][64]
  uint32_t* xhi = tlo + steps * 64;                                  // [steps][kMaxGroups]
  uint32_t* sct = xhi + steps * kMaxGroups;                          // [steps]
  const int* idx = alloc + blockIdx.y * steps;
  for (int w = threadIdx.x; w < steps * 32; w += bd)
    vall[w] = dirs[static_cast<size_t>(idx[w >> 5]) * 32 + (w & 31)];
  for (int w = threadIdx.x; w < steps; w += bd) {
    const int d = idx[w];
    sct[w] = d < number_scrambled ? scr[d] : 0u;
  }
  __syncthreads();
  for (int w = threadIdx.x; w < steps * 64; w += bd) {
    const uint32_t l = static_cast<uint32_t>(w & 63), g = l ^ (l >> 1);
    const uint32_t* vv = vall + (w >> 6) * 32;
    uint32_t x = 0;
#pragma unroll
    for (int bit = 0; bit < 6; ++bit)
      if ((g >> bit) & 1u) x ^= vv[bit];
    tlo[w] = x;
  }
  const int ngroups = bd / 64 + 1;
  const int half = threadIdx.x >> 5, b = threadIdx.x & 31;
  const int halves = bd >> 5, pairs = steps * ngroups;
  const int nchunks = (paths + bd - 1) / bd;
  for (int c = blockIdx.x; c < nchunks; c += gridDim.x) {
    const uint32_t nb = static_cast<uint32_t>(c) * bd + off1;  // n of lane 0
    const uint32_t g0 = nb & ~63u;
    __syncthreads();  // the previous chunk's lanes are done with xhi
    // X(G) for G = g0 + 64k: lane b of a half-wavefront contributes v[b] when
    // bit b of gray(G) is set; the half XOR-reduces.  Every half runs the
    // same number of rounds so the shuffles see all their lanes.
    for (int base = 0; base < pairs; base += halves) {
      const int q = base + half;
      const int qc = q < pairs ? q : pairs - 1;
      const int i = qc / ngroups, k = qc % ngroups;
      const uint32_t G = g0 + 64u * static_cast<uint32_t>(k);
      const uint32_t g = G ^ (G >> 1);
      uint32_t x = ((g >> b) & 1u) ? vall[i * 32 + b] : 0u;
#pragma unroll
      for (int o = 16; o > 0; o >>= 1)
        x ^= static_cast<uint32_t>(__shfl_xor(static_cast<int>(x), o, 32));
      if (b == 0 && q < pairs) xhi[i * kMaxGroups + k] = x ^ sct[i];
    }
    __syncthreads();
    const int p = c * bd + threadIdx.x;
    if (p < paths) {
      const uint32_t n = nb + static_cast<uint32_t>(threadIdx.x);
      const uint32_t kg = (n - g0) >> 6, l = n & 63u;
      bridge_lane(T, steps, factors, paths, p, lds,
                  [&](int i) { return invnorm_u32(xhi[i * kMaxGroups + kg] ^ tlo[i * 64 + l]); },
                  out);
    }
  }
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
  if (steps <= kTableSteps) {
    // LDS: path storage steps x threads floats plus 105 words per step
    // (direction numbers, T(l), X(G), scrambler) within 64 KiB.
    const int threads = steps <= 40 ? 256 : 128;
    const size_t lds = sizeof(float) * steps * threads +
                       sizeof(uint32_t) * steps * (32 + 64 + kMaxGroups + 1);
    // Enough blocks to fill the chip (about four per CU over all factors);
    // each walks several chunks of paths with its tables kept in LDS.
    const int chunks = (paths + threads - 1) / threads;
    int per_factor = (4 * device_cu_count() + factors - 1) / factors;
    if (per_factor > chunks) per_factor = chunks;
    if (per_factor < 1) per_factor = 1;
    hipLaunchKernelGGL(k_sobol_bridge_tab, dim3(per_factor, factors),
                       dim3(threads), lds, as_stream(stream), tables_for(steps), steps, factors,
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
