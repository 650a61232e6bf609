// bridge_device.h -- the Brownian bridge (Jaeckel construction on the unit
// grid) shared by the bridge kernels and the persistent portfolio kernel:
// the construction tables (computed on the host in float, the reference's
// arithmetic) and one lane's bridge built in LDS.
#pragma once

#include <cmath>
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

// One step of the construction, v = rw * right (+ lw * left) + sd * z, with
// the multiply-adds written out so that every form of the bridge (lane
// paths in LDS or in registers, one or four paths per lane) rounds alike and
// gives the same bits.
__device__ __forceinline__ float bridge_fma(float a, float b, float c) {
  return __builtin_fmaf(a, b, c);
}
__device__ __forceinline__ f32x4 bridge_fma(float a, f32x4 b, f32x4 c) {
  return __builtin_elementwise_fma(f32x4{a, a, a, a}, b, c);
}

// Builds one factor's path for this lane's path index from the variate
// source `z(i)` (bridge variate i), differences it and stores (S,F,P).
// f: the factor (the path's bridge of that factor).
template <class Source>
__device__ __forceinline__ void bridge_lane(const BridgeTables& T, int steps, int factors,
                                            int paths, int p, int f, float* lds, Source z,
                                            float* __restrict__ out) {
  const int lane = threadIdx.x;
  float* path = lds + lane;  // path[i * blockDim.x]
  const int bd = blockDim.x;
  path[(steps - 1) * bd] = T.sd[0] * z(0);
  for (int i = 1; i < steps; ++i) {
    const int j = T.left[i], k = T.right[i], l = T.bridge[i];
    float v = T.rw[i] * path[k * bd];
    if (j != 0) v = bridge_fma(T.lw[i], path[(j - 1) * bd], v);
    path[l * bd] = bridge_fma(T.sd[i], z(i), v);
  }
  float prev = 0.0f;
  for (int i = 0; i < steps; ++i) {
    const float w = path[i * bd];
    out[(static_cast<size_t>(i) * factors + f) * paths + p] = w - prev;
    prev = w;
  }
}

}  // namespace
}  // namespace koo
