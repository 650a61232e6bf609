// reductions.hip -- stage (c): path-sum reductions for gfx950.
//
// PartialReduce_using8s/16s keep the reference's pairwise 8- and 16-term
// trees (reductions_gpu.cu:11-118, 218-275) so their outputs match it term
// for term.  The full sums (reduceSinglePass, koo_sum_*) are two-pass and
// deterministic: each block reduces a contiguous chunk with wavefront
// shuffles in FP64, a second single-block pass adds the block partials with
// Neumaier compensation.  No atomics, no retirement counters, so calls are
// re-entrant and capturable.
#include <type_traits>

#include "koo_common.h"

namespace koo {

namespace {

constexpr int kRedThreads = 256;
constexpr int kRedMaxBlocks = 2048;

__global__ void __launch_bounds__(256) k_partial8(const float* __restrict__ in, float* __restrict__ out, int n_out) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n_out; i += gridDim.x * blockDim.x) {
    const float* d = in + static_cast<size_t>(i) * 8;
    float x0 = d[0] + d[1], x2 = d[2] + d[3], x4 = d[4] + d[5], x6 = d[6] + d[7];
    x0 += x2;
    x4 += x6;
    out[i] = x0 + x4;
  }
}

__global__ void __launch_bounds__(256) k_partial16(const float* __restrict__ in, float* __restrict__ out, int n_out) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n_out; i += gridDim.x * blockDim.x) {
    const float* d = in + static_cast<size_t>(i) * 16;
    float x0 = d[0] + d[1], x2 = d[2] + d[3], x4 = d[4] + d[5], x6 = d[6] + d[7];
    x0 += x2;
    x4 += x6;
    x0 += x4;
    float y0 = d[8] + d[9], y2 = d[10] + d[11], y4 = d[12] + d[13], y6 = d[14] + d[15];
    y0 += y2;
    y4 += y6;
    y0 += y4;
    out[i] = x0 + y0;
  }
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Block b sums elements [b*chunk, min((b+1)*chunk, n)) into part[b].
template <class T>
__global__ void __launch_bounds__(kRedThreads)
    k_block_sums(const T* __restrict__ in, long long n, long long chunk, double* __restrict__ part) {
  __shared__ double ws[kRedThreads / 64];
  const long long lo = static_cast<long long>(blockIdx.x) * chunk;
  long long hi = lo + chunk;
  if (hi > n) hi = n;
  double acc = 0.0;
  for (long long i = lo + threadIdx.x; i < hi; i += blockDim.x) acc += static_cast<double>(in[i]);
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < static_cast<int>(blockDim.x >> 6); ++w) t += ws[w];
    part[blockIdx.x] = t;
  }
}

// FP64 sum of the four lanes of a float4.
__device__ __forceinline__ double sum4(const f32x4 x) {
  return (static_cast<double>(x.x) + static_cast<double>(x.y)) +
         (static_cast<double>(x.z) + static_cast<double>(x.w));
}

// The same for FP32 input, 16 bytes per load: block b sums float4 elements
// [b*chunk4, min((b+1)*chunk4, n4)) with four loads in flight per lane
// (streaming, nontemporal) into four FP64 accumulators; the last block also
// adds the n_tail (< 4) elements past the last float4.
__global__ void __launch_bounds__(kRedThreads)
    k_block_sums_vec4(const f32x4* __restrict__ in, long long n4, long long chunk4,
                      const float* __restrict__ tail, int n_tail, double* __restrict__ part) {
  __shared__ double ws[kRedThreads / 64];
  const long long lo = static_cast<long long>(blockIdx.x) * chunk4;
  long long hi = lo + chunk4;
  if (hi > n4) hi = n4;
  const long long B = blockDim.x;
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
  long long i = lo + threadIdx.x;
  for (; i + 3 * B < hi; i += 4 * B) {
    const f32x4 x = load_stream(in + i);
    const f32x4 y = load_stream(in + i + B);
    const f32x4 u = load_stream(in + i + 2 * B);
    const f32x4 v = load_stream(in + i + 3 * B);
    a0 += sum4(x);
    a1 += sum4(y);
    a2 += sum4(u);
    a3 += sum4(v);
  }
  for (; i < hi; i += B) a0 += sum4(load_stream(in + i));
  // the last block's first lane also takes the 0-3 elements past the float4s
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0)
    for (int t = 0; t < n_tail; ++t) a0 += static_cast<double>(tail[t]);
  double acc = wave_sum((a0 + a1) + (a2 + a3));
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < static_cast<int>(blockDim.x >> 6); ++w) t += ws[w];
    part[blockIdx.x] = t;
  }
}

// One block: compensated sum of nb partials; result to out_d and/or out_f.
// Each lane adds a strided slice with Neumaier's correction, then the lanes'
// (sum, correction) pairs are combined by a wavefront butterfly of
// error-free additions (TwoSum) and the four waves' pairs in LDS, so no lane
// walks more than nb / 256 + 6 + 4 dependent additions.
__global__ void __launch_bounds__(kRedThreads)
    k_final_sum(const double* __restrict__ part, int nb, double* __restrict__ out_d,
                float* __restrict__ out_f) {
  __shared__ double sv[kRedThreads / 64], sc[kRedThreads / 64];
  double s = 0.0, c = 0.0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) {
    const double x = part[i];
    const double t = s + x;
    c += (fabs(s) >= fabs(x)) ? (s - t) + x : (x - t) + s;
    s = t;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const double s2 = __shfl_xor(s, o, 64), c2 = __shfl_xor(c, o, 64);
    double t, e;
    two_sum(s, s2, t, e);
    s = t;
    c += c2 + e;
  }
  if ((threadIdx.x & 63) == 0) {
    sv[threadIdx.x >> 6] = s;
    sc[threadIdx.x >> 6] = c;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double S = 0.0, C = 0.0;
    for (int w = 0; w < static_cast<int>(blockDim.x >> 6); ++w) {
      double t, e;
      two_sum(S, sv[w], t, e);
      S = t;
      C += sc[w] + e;
    }
    if (out_d) *out_d = S + C;
    if (out_f) *out_f = static_cast<float>(S + C);
  }
}

int blocks_for_sum(long long n, long long* chunk) {
  long long nb = (n + 4095) / 4096;
  if (nb > kRedMaxBlocks) nb = kRedMaxBlocks;
  if (nb < 1) nb = 1;
  *chunk = (n + nb - 1) / nb;
  return static_cast<int>(nb);
}

template <class T>
int sum_impl(const char* who, const T* in, long long n, double* wsp, double* out_d, float* out_f,
             void* stream) {
  if (n < 0) return set_error(KOO_ERR_INVALID, "%s: n < 0", who);
  if (!wsp || (!out_d && !out_f) || (n > 0 && !in))
    return set_error(KOO_ERR_INVALID, "%s: null pointer", who);
  long long chunk = 0;
  const int nb = blocks_for_sum(n, &chunk);
  hipStream_t st = as_stream(stream);
  if constexpr (std::is_same<T, float>::value) {
    if (n >= 4 && (reinterpret_cast<uintptr_t>(in) & 15u) == 0) {
      const long long n4 = n / 4;
      hipLaunchKernelGGL(k_block_sums_vec4, dim3(nb), dim3(kRedThreads), 0, st,
                         reinterpret_cast<const f32x4*>(in), n4, (n4 + nb - 1) / nb,
                         in + 4 * n4, static_cast<int>(n - 4 * n4), wsp);
      hipLaunchKernelGGL(k_final_sum, dim3(1), dim3(kRedThreads), 0, st, wsp, nb, out_d, out_f);
      return check_launch(who);
    }
  }
  hipLaunchKernelGGL(k_block_sums<T>, dim3(nb), dim3(kRedThreads), 0, st, in, n, chunk, wsp);
  hipLaunchKernelGGL(k_final_sum, dim3(1), dim3(kRedThreads), 0, st, wsp, nb, out_d, out_f);
  return check_launch(who);
}

int partial_reduce(const char* who, int width, const float* data, float* workspace, int points,
                   int stopPoint, float** result, int* remaining, void* stream) {
  if (points < 0 || stopPoint < 1 || !result || !remaining)
    return set_error(KOO_ERR_INVALID, "%s: bad arguments", who);
  if (points > 0 && (!data || !workspace)) return set_error(KOO_ERR_INVALID, "%s: null pointer", who);
  const float* cur = data;
  float* next = workspace;
  int left = points;
  while (left > stopPoint) {
    const int n_out = left / width;
    if (n_out == 0) break;
    if (width == 8)
      hipLaunchKernelGGL(k_partial8, dim3(grid_for(n_out, 256, 4096)), dim3(256), 0,
                         as_stream(stream), cur, next, n_out);
    else
      hipLaunchKernelGGL(k_partial16, dim3(grid_for(n_out, 256, 4096)), dim3(256), 0,
                         as_stream(stream), cur, next, n_out);
    cur = next;
    next += n_out;
    left = n_out;
  }
  *result = const_cast<float*>(cur);
  *remaining = left;
  return check_launch(who);
}

}  // namespace
}  // namespace koo

using namespace koo;

extern "C" int PartialReduce_using8s_gpu(const float* data, float* workspace, int points,
                                         int /*threads*/, int stopPoint, float** result,
                                         int* remaining, void* stream) {
  return partial_reduce("PartialReduce_using8s_gpu", 8, data, workspace, points, stopPoint, result,
                        remaining, stream);
}

extern "C" int PartialReduce_using16s_gpu(const float* data, float* workspace, int points,
                                          int /*threads*/, int stopPoint, float** result,
                                          int* remaining, void* stream) {
  return partial_reduce("PartialReduce_using16s_gpu", 16, data, workspace, points, stopPoint,
                        result, remaining, stream);
}

extern "C" int reduceSinglePass(int size, int /*threads*/, int /*blocks*/, float* in, float* wsp,
                                float* final_loc, void* stream) {
  // wsp holds >= 1024 floats = 512 doubles of block partials; cap the blocks.
  if (size < 0) return set_error(KOO_ERR_INVALID, "reduceSinglePass: size < 0");
  if (!wsp || !final_loc || (size > 0 && !in))
    return set_error(KOO_ERR_INVALID, "reduceSinglePass: null pointer");
  if ((reinterpret_cast<uintptr_t>(wsp) & 7u) != 0)
    return set_error(KOO_ERR_INVALID, "reduceSinglePass: workspace must be 8-byte aligned");
  long long n = size;
  long long nb = (n + 4095) / 4096;
  if (nb > 512) nb = 512;
  if (nb < 1) nb = 1;
  const long long chunk = (n + nb - 1) / nb;
  double* part = reinterpret_cast<double*>(wsp);
  hipStream_t st = as_stream(stream);
  hipLaunchKernelGGL(k_block_sums<float>, dim3(static_cast<unsigned>(nb)), dim3(kRedThreads), 0, st,
                     in, n, chunk, part);
  hipLaunchKernelGGL(k_final_sum, dim3(1), dim3(kRedThreads), 0, st, part, static_cast<int>(nb),
                     static_cast<double*>(nullptr), final_loc);
  return check_launch("reduceSinglePass");
}

extern "C" long long koo_sum_workspace_doubles(long long n) {
  long long chunk = 0;
  return blocks_for_sum(n < 0 ? 0 : n, &chunk);
}

extern "C" int koo_sum_f32_to_f64(const float* in, long long n, double* workspace, double* out,
                                  void* stream) {
  return sum_impl<float>("koo_sum_f32_to_f64", in, n, workspace, out, nullptr, stream);
}

extern "C" int koo_sum_f64(const double* in, long long n, double* workspace, double* out,
                           void* stream) {
  return sum_impl<double>("koo_sum_f64", in, n, workspace, out, nullptr, stream);
}
