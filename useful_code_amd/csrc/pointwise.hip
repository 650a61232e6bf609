// pointwise.hip -- the reference's pointwise products of path-minor rows
// (outerProduct_gpu.h:94-115), the per-path terms whose path sums are the
// outer products.  Pure HBM streams: lane per path (four paths per lane when
// rows are 16-byte aligned), the row loop inside the lane so every load and
// store of a wavefront is one contiguous row segment.
#include "koo_common.h"

namespace koo {
namespace {

constexpr int kPwThreads = 256;

template <class V>
__device__ __forceinline__ V mul(V a, V b);
template <>
__device__ __forceinline__ float mul(float a, float b) {
  return a * b;
}
template <>
__device__ __forceinline__ float4 mul(float4 a, float4 b) {
  return make_float4(a.x * b.x, a.y * b.y, a.z * b.z, a.w * b.w);
}

// out[p] = a[p] b[p]; V = float4 covers four paths (n counts V elements).
template <class V>
__global__ void __launch_bounds__(kPwThreads)
    k_pointwise(const V* __restrict__ a, const V* __restrict__ b, V* __restrict__ out, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    out[i] = mul(a[i], b[i]);
}

// out[r][c][p] = x[r][p] x[c][p] for c >= r (rows x rows x n, lower
// triangle untouched).
template <class V>
__global__ void __launch_bounds__(kPwThreads)
    k_pointwise_sym_upper(const V* __restrict__ x, V* __restrict__ out, int n, int rows) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    for (int r = 0; r < rows; ++r) {
      const V xr = x[static_cast<size_t>(r) * n + i];
      for (int c = r; c < rows; ++c)
        out[(static_cast<size_t>(r) * rows + c) * n + i] =
            mul(xr, c == r ? xr : x[static_cast<size_t>(c) * n + i]);
    }
  }
}

// out[r][p] = a[r][p] b[p].
template <class V>
__global__ void __launch_bounds__(kPwThreads)
    k_pointwise_multiple_first(const V* __restrict__ a, const V* __restrict__ b,
                               V* __restrict__ out, int n, int rows) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const V y = b[i];
    for (int r = 0; r < rows; ++r)
      out[static_cast<size_t>(r) * n + i] = mul(a[static_cast<size_t>(r) * n + i], y);
  }
}

int grid_for(int n) {
  const int cap = 8 * device_cu_count();
  const int g = (n + kPwThreads - 1) / kPwThreads;
  return g < cap ? (g > 0 ? g : 1) : cap;
}

}  // namespace
}  // namespace koo

using namespace koo;

// outerProduct_gpu.h:94-99 (PointwiseProduct_gpu).  blocks / threads are the
// reference's launch shape and are ignored (the grid fills the device).
extern "C" int PointwiseProduct_gpu(int blocks, int threads, int paths, const float* in1_global,
                                    const float* in2_global, float* out_global, void* stream) {
  (void)blocks;
  (void)threads;
  if (paths < 0) return set_error(KOO_ERR_INVALID, "PointwiseProduct_gpu: paths %d", paths);
  if (paths == 0) return KOO_OK;
  if (!in1_global || !in2_global || !out_global)
    return set_error(KOO_ERR_INVALID, "PointwiseProduct_gpu: null pointer");
  hipStream_t st = as_stream(stream);
  if (paths % 4 == 0 && aligned16(in1_global) && aligned16(in2_global) && aligned16(out_global)) {
    const int n = paths / 4;
    hipLaunchKernelGGL(k_pointwise<float4>, dim3(grid_for(n)), dim3(kPwThreads), 0, st,
                       reinterpret_cast<const float4*>(in1_global),
                       reinterpret_cast<const float4*>(in2_global),
                       reinterpret_cast<float4*>(out_global), n);
  } else {
    hipLaunchKernelGGL(k_pointwise<float>, dim3(grid_for(paths)), dim3(kPwThreads), 0, st,
                       in1_global, in2_global, out_global, paths);
  }
  return check_launch("PointwiseProduct_gpu");
}

// outerProduct_gpu.h:101-106 (PointwiseProductSymmetricUpperTriangular_gpu):
// in (rows, paths) -> out (rows, rows, paths), entries c >= r of row r.
extern "C" int PointwiseProductSymmetricUpperTriangular_gpu(int blocks, int threads, int paths,
                                                            int rows, const float* in_global,
                                                            float* out_global, void* stream) {
  (void)blocks;
  (void)threads;
  if (paths < 0 || rows < 0)
    return set_error(KOO_ERR_INVALID,
                     "PointwiseProductSymmetricUpperTriangular_gpu: paths %d rows %d", paths, rows);
  if (paths == 0 || rows == 0) return KOO_OK;
  if (!in_global || !out_global)
    return set_error(KOO_ERR_INVALID, "PointwiseProductSymmetricUpperTriangular_gpu: null pointer");
  hipStream_t st = as_stream(stream);
  if (paths % 4 == 0 && aligned16(in_global) && aligned16(out_global)) {
    const int n = paths / 4;
    hipLaunchKernelGGL(k_pointwise_sym_upper<float4>, dim3(grid_for(n)), dim3(kPwThreads), 0, st,
                       reinterpret_cast<const float4*>(in_global),
                       reinterpret_cast<float4*>(out_global), n, rows);
  } else {
    hipLaunchKernelGGL(k_pointwise_sym_upper<float>, dim3(grid_for(paths)), dim3(kPwThreads), 0,
                       st, in_global, out_global, paths, rows);
  }
  return check_launch("PointwiseProductSymmetricUpperTriangular_gpu");
}

// outerProduct_gpu.h:109-115 (PointwiseProductMultipleFirst_gpu):
// in1 (rows, paths), in2 (paths) -> out (rows, paths).
extern "C" int PointwiseProductMultipleFirst_gpu(int blocks, int threads, int paths,
                                                 const float* in1_global, const float* in2_global,
                                                 float* out_global, int rows, void* stream) {
  (void)blocks;
  (void)threads;
  if (paths < 0 || rows < 0)
    return set_error(KOO_ERR_INVALID, "PointwiseProductMultipleFirst_gpu: paths %d rows %d",
                     paths, rows);
  if (paths == 0 || rows == 0) return KOO_OK;
  if (!in1_global || !in2_global || !out_global)
    return set_error(KOO_ERR_INVALID, "PointwiseProductMultipleFirst_gpu: null pointer");
  hipStream_t st = as_stream(stream);
  if (paths % 4 == 0 && aligned16(in1_global) && aligned16(in2_global) && aligned16(out_global)) {
    const int n = paths / 4;
    hipLaunchKernelGGL(k_pointwise_multiple_first<float4>, dim3(grid_for(n)), dim3(kPwThreads), 0,
                       st, reinterpret_cast<const float4*>(in1_global),
                       reinterpret_cast<const float4*>(in2_global),
                       reinterpret_cast<float4*>(out_global), n, rows);
  } else {
    hipLaunchKernelGGL(k_pointwise_multiple_first<float>, dim3(grid_for(paths)), dim3(kPwThreads),
                       0, st, in1_global, in2_global, out_global, paths, rows);
  }
  return check_launch("PointwiseProductMultipleFirst_gpu");
}
