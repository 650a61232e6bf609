// sobol.hip -- stage (a): Sobol integers, digital-shift scrambling and the
// normal quantile, for gfx950.
//
// Layout: [D][P] (dimension-major, path-minor), the reference's layout
// (sobol_int_gpu.cu:76-161).  Each block owns one dimension (blockIdx.y) and
// a power-of-two stride of points; a lane computes its first point from the
// Gray code and the rest with the two-term stride skip, so every store of a
// wavefront is one contiguous 256-byte segment.
#include <cstdarg>
#include <cstdio>
#include <cstring>

#include "koo_common.h"

// Joe-Kuo (degree, a, m[]) rows, one per dimension (host data); degree <= 17
// (max_m, gold/sobol_primitives.h:67).
struct koo_sobol_primitive {
  unsigned degree;
  unsigned a;
  unsigned m[17];
};
#include "sobol_joe_kuo.inc"

namespace koo {

namespace {
constexpr int kSobolThreads = 256;
constexpr int kPointsPerThread = 8;
// 16-byte kernels: quads of points per lane.  Each block first loads its
// dimension's direction numbers and forms its lanes' first points from the
// Gray code (a loop over the bits); eight quads per lane spread that set-up
// over 128 bytes written per lane.
constexpr int kQuadsPerThread = 8;

__host__ __device__ inline unsigned floor_log2(unsigned x) { return 31u - __builtin_clz(x); }

// Shared body: the integer of point i (Gray index i + off1) for every i this
// lane owns, handed to `emit`.
template <class Emit>
__device__ __forceinline__ void sobol_strided(const uint32_t* v, uint32_t n_vectors, uint32_t off1,
                                              Emit emit) {
  const uint32_t stride = gridDim.x * blockDim.x;  // power of two
  const uint32_t m = floor_log2(stride);
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_vectors) return;
  uint32_t x = sobol_point(v, i + off1);
  emit(i, x);
  const uint32_t v_stride = v[m - 1];
  for (i += stride; i < n_vectors; i += stride) {
    // x(n + 2^m) = x(n) ^ v[m-1] ^ v[c], c = lowest zero bit of n at or above m.
    const uint32_t t = ~((i - stride + off1) | (stride - 1u));
    const uint32_t c = t ? __builtin_ctz(t) : 31u;
    x ^= v_stride ^ v[c];
    emit(i, x);
  }
}

// Four consecutive points per lane: Gray indices n .. n+3 (the Gray codes of
// n and n+1 differ in bit ctz(n+1) only, so each point is the previous one
// XOR one direction number), quads strided by a power of two with the skip
// above at point stride 2^m = 4 x the quad stride.
template <class Emit4>
__device__ __forceinline__ void sobol_strided4(const uint32_t* v, uint32_t n_quads, uint32_t off1,
                                               Emit4 emit) {
  const uint32_t stride = gridDim.x * blockDim.x;  // quads, a power of two
  const uint32_t m = floor_log2(stride) + 2u;
  uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n_quads) return;
  uint32_t n = 4u * q + off1;
  uint32_t x = sobol_point(v, n);
  const uint32_t v_stride = v[m - 1];
  for (;;) {
    const uint32_t x1 = x ^ v[__builtin_ctz(n + 1u)];
    const uint32_t x2 = x1 ^ v[__builtin_ctz(n + 2u)];
    const uint32_t x3 = x2 ^ v[__builtin_ctz(n + 3u)];
    emit(q, x, x1, x2, x3);
    q += stride;
    if (q >= n_quads) break;
    const uint32_t t = ~(n | ((1u << m) - 1u));
    x ^= v_stride ^ v[t ? __builtin_ctz(t) : 31u];
    n += 1u << m;
  }
}

__global__ void __launch_bounds__(kSobolThreads) k_sobol_ints_vec4(uint32_t n_quads, uint32_t off1,
                                                                   const uint32_t* __restrict__ dirs,
                                                                   u32x4* __restrict__ out) {
  __shared__ uint32_t v[32];
  if (threadIdx.x < 32) v[threadIdx.x] = dirs[32u * blockIdx.y + threadIdx.x];
  __syncthreads();
  u32x4* o = out + static_cast<size_t>(blockIdx.y) * n_quads;
  sobol_strided4(v, n_quads, off1, [&](uint32_t q, uint32_t a, uint32_t b, uint32_t c, uint32_t d) {
    const u32x4 w = {a, b, c, d};
    store_stream(w, o + q);
  });
}

__global__ void __launch_bounds__(kSobolThreads)
    k_sobol_normal_vec4(uint32_t n_quads, uint32_t off1, const uint32_t* __restrict__ dirs,
                        const uint32_t* __restrict__ scramblers, int number_scrambled,
                        f32x4* __restrict__ out) {
  __shared__ uint32_t v[32];
  if (threadIdx.x < 32) v[threadIdx.x] = dirs[32u * blockIdx.y + threadIdx.x];
  __syncthreads();
  const uint32_t s = (static_cast<int>(blockIdx.y) < number_scrambled) ? scramblers[blockIdx.y] : 0u;
  f32x4* o = out + static_cast<size_t>(blockIdx.y) * n_quads;
  sobol_strided4(v, n_quads, off1, [&](uint32_t q, uint32_t a, uint32_t b, uint32_t c, uint32_t d) {
    const f32x4 w = {invnorm_u32(a ^ s), invnorm_u32(b ^ s), invnorm_u32(c ^ s),
                     invnorm_u32(d ^ s)};
    store_stream(w, o + q);
  });
}

__global__ void __launch_bounds__(kSobolThreads) k_sobol_ints(uint32_t n_vectors, uint32_t off1,
                                                              const uint32_t* __restrict__ dirs,
                                                              uint32_t* __restrict__ out) {
  __shared__ uint32_t v[32];
  if (threadIdx.x < 32) v[threadIdx.x] = dirs[32u * blockIdx.y + threadIdx.x];
  __syncthreads();
  uint32_t* o = out + static_cast<size_t>(blockIdx.y) * n_vectors;
  sobol_strided(v, n_vectors, off1, [&](uint32_t i, uint32_t x) { o[i] = x; });
}

__global__ void __launch_bounds__(kSobolThreads)
    k_sobol_normal(uint32_t n_vectors, uint32_t off1, const uint32_t* __restrict__ dirs,
                   const uint32_t* __restrict__ scramblers, int number_scrambled,
                   float* __restrict__ out) {
  __shared__ uint32_t v[32];
  if (threadIdx.x < 32) v[threadIdx.x] = dirs[32u * blockIdx.y + threadIdx.x];
  __syncthreads();
  const uint32_t s = (static_cast<int>(blockIdx.y) < number_scrambled) ? scramblers[blockIdx.y] : 0u;
  float* o = out + static_cast<size_t>(blockIdx.y) * n_vectors;
  sobol_strided(v, n_vectors, off1, [&](uint32_t i, uint32_t x) { o[i] = invnorm_u32(x ^ s); });
}

// In-place XOR of the first number_dim rows; 4 words per lane when rows are
// 16-byte aligned.  One row per blockIdx.y, so the scrambler is one scalar
// load per block and no lane divides to find its row.
__global__ void k_scramble_rows(u32x4* __restrict__ data, const uint32_t* __restrict__ scr,
                                int paths4) {
  const uint32_t s = scr[blockIdx.y];
  u32x4* row = data + static_cast<size_t>(blockIdx.y) * paths4;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < paths4; i += gridDim.x * blockDim.x)
    store_stream(load_stream(row + i) ^ s, row + i);
}

__global__ void k_scramble(uint32_t* __restrict__ data, const uint32_t* __restrict__ scr,
                           int64_t n, int64_t paths) {
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)
    data[i] ^= scr[i / paths];
}

__global__ void k_invnorm_vec4(const u32x4* __restrict__ in, f32x4* __restrict__ out, int64_t n4) {
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n4;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const u32x4 w = load_stream(in + i);
    const f32x4 z = {invnorm_u32(w.x), invnorm_u32(w.y), invnorm_u32(w.z), invnorm_u32(w.w)};
    store_stream(z, out + i);
  }
}

__global__ void k_invnorm(const uint32_t* __restrict__ in, float* __restrict__ out, int64_t begin,
                          int64_t n) {
  for (int64_t i = begin + blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)
    out[i] = invnorm_u32(in[i]);
}

unsigned next_pow2(unsigned x) {
  unsigned p = 1;
  while (p < x) p <<= 1;
  return p;
}

dim3 sobol_grid4(int n_quads, int n_dimensions) {
  unsigned per_block = kSobolThreads * kQuadsPerThread;
  unsigned gx = next_pow2((static_cast<unsigned>(n_quads) + per_block - 1) / per_block);
  return dim3(gx, static_cast<unsigned>(n_dimensions));
}

dim3 sobol_grid(int n_vectors, int n_dimensions) {
  unsigned per_block = kSobolThreads * kPointsPerThread;
  unsigned gx = next_pow2((static_cast<unsigned>(n_vectors) + per_block - 1) / per_block);
  return dim3(gx, static_cast<unsigned>(n_dimensions));
}

}  // namespace

// ---------------------------------------------------------------- errors

static thread_local char g_last_error[512] = "";

int set_error(int status, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
  va_end(ap);
  return status;
}

int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error(KOO_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
  return KOO_OK;
}

int device_cu_count() {
  static int cached[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  if (!cached[dev]) {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus <= 0)
      cus = 256;
    cached[dev] = cus;
  }
  return cached[dev];
}

}  // namespace koo

using namespace koo;

extern "C" const char* koo_last_error(void) { return g_last_error; }
extern "C" const char* koo_version(void) { return "useful_code_amd 0.1 (gfx950)"; }

// Bratley-Fox direction numbers; the recurrence of gold/sobol_gold.cpp:118-164.
extern "C" int initSobolDirectionVectors(int n_dimensions, unsigned int* directions) {
  if (n_dimensions < 1 || n_dimensions > KOO_SOBOL_TABLE_DIMS || !directions)
    return set_error(KOO_ERR_INVALID, "initSobolDirectionVectors: n_dimensions %d outside [1,%d]",
                     n_dimensions, KOO_SOBOL_TABLE_DIMS);
  for (int d = 0; d < n_dimensions; ++d) {
    unsigned* v = directions + 32 * d;
    if (d == 0) {
      for (int i = 0; i < 32; ++i) v[i] = 1u << (31 - i);
      continue;
    }
    const int deg = static_cast<int>(koo_sobol_table[d].degree);
    const unsigned a = koo_sobol_table[d].a;
    for (int i = 0; i < deg; ++i) v[i] = koo_sobol_table[d].m[i] << (31 - i);
    for (int i = deg; i < 32; ++i) {
      unsigned x = v[i - deg] ^ (v[i - deg] >> deg);
      for (int j = 1; j < deg; ++j) x ^= ((a >> (deg - 1 - j)) & 1u) * v[i - j];
      v[i] = x;
    }
  }
  return KOO_OK;
}

extern "C" int sobolGPUInts(int n_vectors, int n_dimensions, unsigned path_offset,
                            unsigned int* d_directions, unsigned int* d_output, void* stream) {
  if (n_vectors < 0 || n_dimensions < 0 || n_dimensions > 65535)
    return set_error(KOO_ERR_INVALID, "sobolGPUInts: bad sizes (%d, %d)", n_vectors, n_dimensions);
  if (n_vectors == 0 || n_dimensions == 0) return KOO_OK;
  if (!d_directions || !d_output) return set_error(KOO_ERR_INVALID, "sobolGPUInts: null pointer");
  if (n_vectors % 4 == 0 && aligned16(d_output))
    hipLaunchKernelGGL(k_sobol_ints_vec4, sobol_grid4(n_vectors / 4, n_dimensions),
                       dim3(kSobolThreads), 0, as_stream(stream),
                       static_cast<uint32_t>(n_vectors / 4), path_offset + 1u, d_directions,
                       reinterpret_cast<u32x4*>(d_output));
  else
    hipLaunchKernelGGL(k_sobol_ints, sobol_grid(n_vectors, n_dimensions), dim3(kSobolThreads), 0,
                       as_stream(stream), static_cast<uint32_t>(n_vectors), path_offset + 1u,
                       d_directions, d_output);
  return check_launch("sobolGPUInts");
}

extern "C" int sobol_scramble_normal_gpu(int n_vectors, int n_dimensions, unsigned path_offset,
                                         const unsigned int* d_directions,
                                         const unsigned int* scramblers, int number_scrambled,
                                         float* out, void* stream) {
  if (n_vectors < 0 || n_dimensions < 0 || n_dimensions > 65535)
    return set_error(KOO_ERR_INVALID, "sobol_scramble_normal_gpu: bad sizes");
  if (n_vectors == 0 || n_dimensions == 0) return KOO_OK;
  if (!d_directions || !out || (number_scrambled > 0 && !scramblers))
    return set_error(KOO_ERR_INVALID, "sobol_scramble_normal_gpu: null pointer");
  if (n_vectors % 4 == 0 && aligned16(out))
    hipLaunchKernelGGL(k_sobol_normal_vec4, sobol_grid4(n_vectors / 4, n_dimensions),
                       dim3(kSobolThreads), 0, as_stream(stream),
                       static_cast<uint32_t>(n_vectors / 4), path_offset + 1u, d_directions,
                       scramblers, number_scrambled, reinterpret_cast<f32x4*>(out));
  else
    hipLaunchKernelGGL(k_sobol_normal, sobol_grid(n_vectors, n_dimensions), dim3(kSobolThreads),
                       0, as_stream(stream), static_cast<uint32_t>(n_vectors), path_offset + 1u,
                       d_directions, scramblers, number_scrambled, out);
  return check_launch("sobol_scramble_normal_gpu");
}

extern "C" int scramble_gpu(unsigned int* device_data, unsigned int* scramblers, int number_dim,
                            int paths, void* stream) {
  if (number_dim < 0 || paths < 0) return set_error(KOO_ERR_INVALID, "scramble_gpu: bad sizes");
  const int64_t n = static_cast<int64_t>(number_dim) * paths;
  if (n == 0) return KOO_OK;
  if (!device_data || !scramblers) return set_error(KOO_ERR_INVALID, "scramble_gpu: null pointer");
  const int cap = 8 * device_cu_count() * 8;
  if (paths % 4 == 0 && (reinterpret_cast<uintptr_t>(device_data) & 15u) == 0 &&
      number_dim <= 65535) {
    const int paths4 = paths / 4;
    hipLaunchKernelGGL(k_scramble_rows, dim3(grid_for(paths4, 256, 65535), number_dim), dim3(256),
                       0, as_stream(stream), reinterpret_cast<u32x4*>(device_data), scramblers,
                       paths4);
  } else {
    hipLaunchKernelGGL(k_scramble, dim3(grid_for(n, 256, cap)), dim3(256), 0, as_stream(stream),
                       device_data, scramblers, n, static_cast<int64_t>(paths));
  }
  return check_launch("scramble_gpu");
}

extern "C" int inverse_normal_u32_gpu(const unsigned int* in, float* out, long long n,
                                      void* stream) {
  if (n < 0) return set_error(KOO_ERR_INVALID, "inverse_normal_u32_gpu: n < 0");
  if (n == 0) return KOO_OK;
  if (!in || !out) return set_error(KOO_ERR_INVALID, "inverse_normal_u32_gpu: null pointer");
  const int cap = 8 * device_cu_count() * 8;
  int64_t done = 0;
  if ((reinterpret_cast<uintptr_t>(in) & 15u) == 0 && (reinterpret_cast<uintptr_t>(out) & 15u) == 0) {
    const int64_t n4 = n / 4;
    if (n4 > 0)
      hipLaunchKernelGGL(k_invnorm_vec4, dim3(grid_for(n4, 256, cap)), dim3(256), 0,
                         as_stream(stream), reinterpret_cast<const u32x4*>(in),
                         reinterpret_cast<f32x4*>(out), n4);
    done = n4 * 4;
  }
  if (done < n)
    hipLaunchKernelGGL(k_invnorm, dim3(grid_for(n - done, 256, cap)), dim3(256), 0,
                       as_stream(stream), in, out, done, static_cast<int64_t>(n));
  return check_launch("inverse_normal_u32_gpu");
}
