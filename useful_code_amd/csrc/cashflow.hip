// cashflow.hip -- stage (c): product cash-flow generation, discounting to the
// numeraire and aggregation onto exercise dates, for gfx950.
//
// One lane per path; product constants and per-step index/weight tables are
// read with uniform addresses (scalar unit), per-path data is path-minor so
// every wavefront access is one contiguous row.
#include "koo_common.h"

namespace koo {

namespace {

// Product state machines: getCashFlows(flow1, flow2, rate1, rate2, rate3)
// returns true once the product has terminated on this path.

// Swap of rate1 against a fixed coupon; aux = [K, payReceive, acc1[S], acc2[S]]
// (cashFlowGeneration_product_gpu.h:128-188).
struct SwapProduct {
  const float* aux;
  int steps, index;
  float strike, pay_receive;
  __device__ SwapProduct(const float* a, int s)
      : aux(a), steps(s), index(0), strike(a[0]), pay_receive(a[1]) {}
  __device__ bool flows(float& f1, float& f2, float r1, float, float) {
    f2 = aux[2 + steps + index] * r1 * pay_receive;
    f1 = -strike * aux[2 + index] * pay_receive;
    ++index;
    return index == steps;
  }
};

// Target redemption note; aux = [total coupon, K, m, acc1[S], acc2[S]]
// (cashFlowGeneration_product_gpu.h:19-108).
struct TarnProduct {
  const float* aux;
  int steps, index;
  float total, strike, mult, paid;
  __device__ TarnProduct(const float* a, int s)
      : aux(a), steps(s), index(0), total(a[0]), strike(a[1]), mult(a[2]), paid(0.0f) {}
  __device__ bool flows(float& f1, float& f2, float r1, float r2, float) {
    f2 = aux[3 + steps + index] * r2;
    if (index + 1 == steps) {
      f1 = paid - total;
      return true;
    }
    float unc = (strike - mult * r1) * aux[3 + index];
    if (unc < 0.0f) unc = 0.0f;
    const float nt = paid + unc;
    if (nt >= total) {
      f1 = paid - total;
      return true;
    }
    f1 = -unc;
    paid = nt;
    ++index;
    return false;
  }
};

// Caplet strip on rate1; aux = [K, acc[S]].  No reference counterpart.
struct CapletProduct {
  const float* aux;
  int steps, index;
  float strike;
  __device__ CapletProduct(const float* a, int s) : aux(a), steps(s), index(0), strike(a[0]) {}
  __device__ bool flows(float& f1, float& f2, float r1, float, float) {
    f1 = aux[1 + index] * fmaxf(r1 - strike, 0.0f);
    f2 = 0.0f;
    ++index;
    return index == steps;
  }
};

// Driver loop of cashFlowGenerator_kernel (cashFlowGeneration_gpu.h:23-77).
template <class Product>
__global__ void __launch_bounds__(256)
    k_cashflows(float* __restrict__ f1, float* __restrict__ f2, const float* __restrict__ aux,
                int paths, int steps, const float* __restrict__ r1, const float* __restrict__ r2,
                const float* __restrict__ r3) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= paths) return;
  Product obj(aux, steps);
  bool done = false;
  size_t o = p;
  for (int i = 0; i < steps; ++i, o += paths) {
    float a = 0.0f, b = 0.0f;
    if (!done) done = obj.flows(a, b, r1[o], r2 ? r2[o] : 0.0f, r3 ? r3[o] : 0.0f);
    f1[o] = a;
    f2[o] = b;
  }
}

// Four paths per lane (16-byte streaming loads and stores): one product
// state machine per path, the same steps as k_cashflows.
template <class Product>
__global__ void __launch_bounds__(256)
    k_cashflows_vec4(f32x4* __restrict__ f1, f32x4* __restrict__ f2, const float* __restrict__ aux,
                     int paths4, int steps, const f32x4* __restrict__ r1,
                     const f32x4* __restrict__ r2, const f32x4* __restrict__ r3) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= paths4) return;
  Product obj[4] = {Product(aux, steps), Product(aux, steps), Product(aux, steps),
                    Product(aux, steps)};
  bool done[4] = {false, false, false, false};
  const f32x4 zero = {0.0f, 0.0f, 0.0f, 0.0f};
  size_t o = p;
  for (int i = 0; i < steps; ++i, o += paths4) {
    const f32x4 x1 = load_stream(r1 + o);
    const f32x4 x2 = r2 ? load_stream(r2 + o) : zero;
    const f32x4 x3 = r3 ? load_stream(r3 + o) : zero;
    f32x4 a = zero, b = zero;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      float ac = 0.0f, bc = 0.0f;
      if (!done[c]) done[c] = obj[c].flows(ac, bc, x1[c], x2[c], x3[c]);
      a[c] = ac;
      b[c] = bc;
    }
    store_stream(a, f1 + o);
    store_stream(b, f2 + o);
  }
}

// (r1)^(1 - theta) (r2)^theta, the log-linear interpolation of the discount
// factor between two rate times.  theta is uniform over a launch's lanes:
// the end points (0: payment on a rate time, the usual case) are exact, in
// between 2^((1 - theta) log2 r1 + theta log2 r2) with the hardware log2 /
// exp2 (v_log_f32 / v_exp_f32, about 1 ulp each; the ratios are positive
// normal numbers) instead of two powf calls.
__device__ __forceinline__ float interp_df(float r1, float r2, float th) {
  if (th == 0.0f) return r1;
  if (th == 1.0f) return r2;
  return __builtin_amdgcn_exp2f((1.0f - th) * __builtin_amdgcn_logf(r1) +
                                th * __builtin_amdgcn_logf(r2));
}

// df = (P1/P_s)^(1-theta) (P2/P_s)^theta / B_s plus per-path sum
// (cashFlowDiscounting_gpu.cu:170-258).  The step loop is unrolled so the
// gathers of several steps are in flight; quotients by the hardware
// reciprocal (about 1 ulp).
__global__ void __launch_bounds__(256)
    k_discounting(const int* __restrict__ first, const int* __restrict__ second,
                  const float* __restrict__ theta, const float* __restrict__ disc,
                  const float* __restrict__ flows, const float* __restrict__ num, int paths,
                  int steps, float* __restrict__ out, float* __restrict__ summed) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= paths) return;
  const size_t P = static_cast<size_t>(paths);
  float total = 0.0f;
#pragma unroll 4
  for (int s = 0; s < steps; ++s) {
    const float* dcol = disc + static_cast<size_t>(s) * (steps + 1) * P + p;
    const float inv_now = __builtin_amdgcn_rcpf(dcol[static_cast<size_t>(s) * P]);
    const float r1 = dcol[static_cast<size_t>(first[s]) * P] * inv_now;
    const float r2 = dcol[static_cast<size_t>(second[s]) * P] * inv_now;
    const float rdf = interp_df(r1, r2, theta[s]);
    const size_t o = static_cast<size_t>(s) * P + p;
    const float fl = flows[o] * (rdf * __builtin_amdgcn_rcpf(num[o]));
    total += fl;
    out[o] = fl;
  }
  summed[p] = total;
}

// The same, four paths per lane with 16-byte streaming loads and stores;
// per path the same operations in the same order.
__global__ void __launch_bounds__(256)
    k_discounting_vec4(const int* __restrict__ first, const int* __restrict__ second,
                       const float* __restrict__ theta, const f32x4* __restrict__ disc,
                       const f32x4* __restrict__ flows, const f32x4* __restrict__ num, int paths4,
                       int steps, f32x4* __restrict__ out, f32x4* __restrict__ summed) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= paths4) return;
  const size_t P = static_cast<size_t>(paths4);
  f32x4 total = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll 2
  for (int s = 0; s < steps; ++s) {
    const f32x4* dcol = disc + static_cast<size_t>(s) * (steps + 1) * P + p;
    const f32x4 now = load_stream(dcol + static_cast<size_t>(s) * P);
    const f32x4 d1 = load_stream(dcol + static_cast<size_t>(first[s]) * P);
    const f32x4 d2 = load_stream(dcol + static_cast<size_t>(second[s]) * P);
    const size_t o = static_cast<size_t>(s) * P + p;
    const f32x4 g = load_stream(flows + o), b = load_stream(num + o);
    const float th = theta[s];
    f32x4 fl;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const float inv_now = __builtin_amdgcn_rcpf(now[c]);
      const float rdf = interp_df(d1[c] * inv_now, d2[c] * inv_now, th);
      fl[c] = g[c] * (rdf * __builtin_amdgcn_rcpf(b[c]));
    }
    total += fl;
    store_stream(fl, out + o);
  }
  store_stream(total, summed + p);
}

// Partial discounting to targetIndices via the numeraire
// (cashFlowDiscounting_gpu.cu:369-440; like that kernel, the discount factor
// is also divided by the step's own ratio P(s, s), which the discount-ratio
// cubes hold as 1).
__global__ void __launch_bounds__(256)
    k_discounting_partial(int n_flows, const int* __restrict__ gen, const int* __restrict__ first,
                          const int* __restrict__ second, const float* __restrict__ theta,
                          const int* __restrict__ target, const float* __restrict__ disc,
                          const float* __restrict__ flows, const float* __restrict__ num,
                          int paths, int steps, float* __restrict__ out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= paths) return;
  const size_t P = static_cast<size_t>(paths);
#pragma unroll 4
  for (int k = 0; k < n_flows; ++k) {
    const int i = gen[k];
    const float* dcol = disc + static_cast<size_t>(i) * (steps + 1) * P + p;
    const float inv_now = __builtin_amdgcn_rcpf(dcol[static_cast<size_t>(i) * P]);
    const float r1 = dcol[static_cast<size_t>(first[k]) * P] * inv_now;
    const float r2 = dcol[static_cast<size_t>(second[k]) * P] * inv_now;
    const float rdf = interp_df(r1, r2, theta[k]);
    float fl = flows[static_cast<size_t>(k) * P + p] * (rdf * inv_now);
    fl *= num[static_cast<size_t>(target[k]) * P + p] *
          __builtin_amdgcn_rcpf(num[static_cast<size_t>(i) * P + p]);
    out[static_cast<size_t>(k) * P + p] = fl;
  }
}

// aggregated(j, offset + p) += flows(i, p), j = preceding[i]
// (cashFlowAggregation_gpu.cu:22-53); negative indices (flows before the
// first exercise date) are skipped as the gold twin does.
__global__ void __launch_bounds__(256)
    k_aggregate(float* __restrict__ agg, int total_paths, int n_ex,
                const float* __restrict__ flows, int batch, int offset, int steps,
                const int* __restrict__ preceding, int start) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= batch) return;
  for (int i = start; i < steps; ++i) {
    const int j = preceding[i];
    if (j < 0 || j >= n_ex) continue;
    agg[static_cast<size_t>(j) * total_paths + p + offset] += flows[static_cast<size_t>(i) * batch + p];
  }
}

// ---------------------------------------------------------------------------
// Cancellable swap under a Longstaff-Schwartz multi-regression strategy: the
// reference's second pass cashFlowGeneratorEE_kernel<Swap, earlyExerciseNull,
// LSAMultiExerciseStrategyQuadratic(Cross)_gpu> (cashFlowGen_earlyExProduct_gpu.h:
// 28-172, strategies :756-1049 and the quadratic one after it).  Per path and
// step: at an exercise date the strategy's estimated continuation value (an
// Andersen shift plus the regression of the date's normalised variables,
// refined through up to `depth` regressions while the net value stays inside
// the date's cut band) against the exercise value (zero: cancellation);
// exercise ends the path's flows, otherwise the swap pays.  The strategy
// tables are the reference's packed arrays; every lane of a wavefront is at
// the same exercise date and depth, so they are read with uniform addresses
// (scalar loads) instead of being staged in shared memory.
//   int data: depth, max variables, row size, variables per date [E]
//   float data: Andersen shifts [E], variable shifts [E][depth][maxv],
//     divisors [E][depth][maxv], lower cuts [E][depth], upper cuts [E][depth],
//     weights [E][depth][row size]
// ---------------------------------------------------------------------------
template <bool CROSS>
__device__ __forceinline__ bool ls_multi_strategy(const float* __restrict__ fd,
                                                  const int* __restrict__ id, int E, int ex,
                                                  const float* __restrict__ sv, int sv_max,
                                                  size_t P, int p, float ex_value, float& value) {
  const int depth_max = id[0], maxv = id[1], row = id[2], nv = id[3 + ex];
  const float* shifts = fd + E + static_cast<size_t>(ex) * depth_max * maxv;
  const float* divs = shifts + static_cast<size_t>(E) * depth_max * maxv;
  const float* lower = fd + E + 2 * static_cast<size_t>(E) * depth_max * maxv +
                       static_cast<size_t>(ex) * depth_max;
  const float* upper = lower + static_cast<size_t>(E) * depth_max;
  const float* weights0 = fd + E + 2 * static_cast<size_t>(E) * depth_max * maxv +
                          2 * static_cast<size_t>(E) * depth_max +
                          static_cast<size_t>(ex) * depth_max * row;
  const float* vars = sv + static_cast<size_t>(ex) * P * sv_max + p;
  for (int depth = 0;;) {
    value = fd[ex];
    const float* w = weights0 + depth * row;
    const float* sh = shifts + maxv * depth;
    const float* dv = divs + maxv * depth;
    value += *w++;
    auto var = [&](int i) { return (vars[static_cast<size_t>(i) * P] - sh[i]) / dv[i]; };
    if (!CROSS) {
      for (int i = 0; i < nv; ++i) {
        const float v = var(i);
        value += (*w++) * v;
        value += (*w++) * v * v;
      }
    } else if (nv == 1) {
      const float v = var(0);
      value += (*w++) * v;
      value += (*w++) * v * v;
    } else if (nv == 2) {
      const float x = var(0), y = var(1);
      value += (*w++) * x;
      value += (*w++) * y;
      value += (*w++) * x * x;
      value += (*w++) * x * y;
      value += (*w++) * y * y;
    } else {
      for (int i = 0; i < nv; ++i) value += (*w++) * var(i);
      for (int i = 0; i < nv; ++i) value += (*w++) * var(i) * var((i + 1) % nv);
      for (int i = 0; i < nv; ++i) {
        const float v = var(i);
        value += (*w++) * v * v;
      }
    }
    ++depth;
    if (depth == depth_max) break;
    if (value - ex_value < lower[depth - 1]) {
      value = lower[depth - 1];
      return true;
    }
    if (value - ex_value > upper[depth - 1]) {
      value = upper[depth - 1];
      return false;
    }
  }
  return value < ex_value;
}

template <bool CROSS>
__global__ void __launch_bounds__(256)
    k_cashflows_ee(float* __restrict__ f1, float* __restrict__ f2, const float* __restrict__ aux,
                   const float* __restrict__ fd, const int* __restrict__ id, int paths,
                   int steps, int n_ex, const unsigned char* __restrict__ is_ex,
                   const float* __restrict__ r1, const float* __restrict__ r2,
                   const float* __restrict__ r3, const float* __restrict__ sv, int sv_max,
                   float* __restrict__ est, int* __restrict__ ex_index) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= paths) return;
  const size_t P = static_cast<size_t>(paths);
  SwapProduct swap(aux, steps);
  bool done = false;
  int i = 0, ex = 0;
  for (; i < steps && !done; ++i) {
    const size_t o = static_cast<size_t>(i) * P + p;
    float a = 0.0f, b = 0.0f;
    const float x1 = r1[o], x2 = r2 ? r2[o] : 0.0f, x3 = r3 ? r3[o] : 0.0f;
    bool exercise = false;
    if (is_ex[i]) {
      const float ev = 0.0f;  // earlyExerciseNull: cancellation pays nothing
      float cont = -1.0e20f;
      exercise = ls_multi_strategy<CROSS>(fd, id, n_ex, ex, sv, sv_max, P, p, ev, cont);
      est[o] = cont;
      if (exercise) {
        a = ev;
        done = true;
      }
      ++ex;
    }
    if (!exercise) done = swap.flows(a, b, x1, x2, x3);
    f1[o] = a;
    f2[o] = b;
  }
  ex_index[p] = i;
  for (; i < steps; ++i) {
    const size_t o = static_cast<size_t>(i) * P + p;
    f1[o] = 0.0f;
    f2[o] = 0.0f;
  }
}

template <class Product>
int launch_cashflows(const char* who, float* f1, float* f2, float* aux, int paths, int steps,
                     float* r1, float* r2, float* r3, void* stream) {
  if (paths < 0 || steps < 0) return set_error(KOO_ERR_INVALID, "%s: bad sizes", who);
  if (paths == 0 || steps == 0) return KOO_OK;
  if (!f1 || !f2 || !aux || !r1) return set_error(KOO_ERR_INVALID, "%s: null pointer", who);
  if (paths % 4 == 0 && aligned16(f1) && aligned16(f2) && aligned16(r1) && aligned16(r2) &&
      aligned16(r3)) {
    const int paths4 = paths / 4;
    hipLaunchKernelGGL(k_cashflows_vec4<Product>, dim3((paths4 + 255) / 256), dim3(256), 0,
                       as_stream(stream), reinterpret_cast<f32x4*>(f1),
                       reinterpret_cast<f32x4*>(f2), aux, paths4, steps,
                       reinterpret_cast<const f32x4*>(r1), reinterpret_cast<const f32x4*>(r2),
                       reinterpret_cast<const f32x4*>(r3));
    return check_launch(who);
  }
  hipLaunchKernelGGL(k_cashflows<Product>, dim3((paths + 255) / 256), dim3(256), 0,
                     as_stream(stream), f1, f2, aux, paths, steps, r1, r2, r3);
  return check_launch(who);
}

}  // namespace
}  // namespace koo

using namespace koo;

extern "C" int cashFlowGeneratorCallerSwap(float* genFlows1, float* genFlows2, float* aux_data,
                                           int paths, int numberSteps, float* rates1,
                                           float* rates2, float* rates3, float* /*forwards*/,
                                           float* /*discRatios*/, void* stream) {
  return launch_cashflows<SwapProduct>("cashFlowGeneratorCallerSwap", genFlows1, genFlows2,
                                       aux_data, paths, numberSteps, rates1, rates2, rates3,
                                       stream);
}

extern "C" int cashFlowGeneratorCallerTARN(float* genFlows1, float* genFlows2, float* aux_data,
                                           int paths, int numberSteps, float* rates1,
                                           float* rates2, float* rates3, float* /*forwards*/,
                                           float* /*discRatios*/, void* stream) {
  if (!rates2) return set_error(KOO_ERR_INVALID, "cashFlowGeneratorCallerTARN: rates2 is null");
  return launch_cashflows<TarnProduct>("cashFlowGeneratorCallerTARN", genFlows1, genFlows2,
                                       aux_data, paths, numberSteps, rates1, rates2, rates3,
                                       stream);
}

extern "C" int cashFlowGeneratorCallerCaplet(float* genFlows1, float* genFlows2, float* aux_data,
                                             int paths, int numberSteps, float* rates1,
                                             float* rates2, float* rates3, float* /*forwards*/,
                                             float* /*discRatios*/, void* stream) {
  return launch_cashflows<CapletProduct>("cashFlowGeneratorCallerCaplet", genFlows1, genFlows2,
                                         aux_data, paths, numberSteps, rates1, rates2, rates3,
                                         stream);
}

extern "C" int cashFlowDiscounting_gpu(int* firstIndex, int* secondIndex, float* theta,
                                       float* discountRatios, float* genFlows,
                                       float* numeraireValues, int paths, int numberSteps,
                                       int /*useTextures*/, int /*useShared*/,
                                       float* discountedFlows, float* summedDiscountedFlows,
                                       void* stream) {
  if (paths < 0 || numberSteps < 0) return set_error(KOO_ERR_INVALID, "cashFlowDiscounting_gpu: sizes");
  if (paths == 0 || numberSteps == 0) return KOO_OK;
  if (!firstIndex || !secondIndex || !theta || !discountRatios || !genFlows || !numeraireValues ||
      !discountedFlows || !summedDiscountedFlows)
    return set_error(KOO_ERR_INVALID, "cashFlowDiscounting_gpu: null pointer");
  if (paths % 4 == 0 && aligned16(discountRatios) && aligned16(genFlows) &&
      aligned16(numeraireValues) && aligned16(discountedFlows) &&
      aligned16(summedDiscountedFlows)) {
    const int paths4 = paths / 4;
    hipLaunchKernelGGL(k_discounting_vec4, dim3((paths4 + 255) / 256), dim3(256), 0,
                       as_stream(stream), firstIndex, secondIndex, theta,
                       reinterpret_cast<const f32x4*>(discountRatios),
                       reinterpret_cast<const f32x4*>(genFlows),
                       reinterpret_cast<const f32x4*>(numeraireValues), paths4, numberSteps,
                       reinterpret_cast<f32x4*>(discountedFlows),
                       reinterpret_cast<f32x4*>(summedDiscountedFlows));
    return check_launch("cashFlowDiscounting_gpu");
  }
  hipLaunchKernelGGL(k_discounting, dim3((paths + 255) / 256), dim3(256), 0, as_stream(stream),
                     firstIndex, secondIndex, theta, discountRatios, genFlows, numeraireValues,
                     paths, numberSteps, discountedFlows, summedDiscountedFlows);
  return check_launch("cashFlowDiscounting_gpu");
}

extern "C" int cashFlowDiscounting_partial_gpu(int numberFlowsPerPath, int* genIndices,
                                               int* firstIndex, int* secondIndex, float* theta,
                                               int* targetIndices, float* discountRatios,
                                               float* genFlows, float* numeraireValues, int paths,
                                               int numberSteps, float* discountedFlows,
                                               void* stream) {
  if (paths < 0 || numberSteps < 0 || numberFlowsPerPath < 0)
    return set_error(KOO_ERR_INVALID, "cashFlowDiscounting_partial_gpu: sizes");
  if (paths == 0 || numberFlowsPerPath == 0) return KOO_OK;
  if (!genIndices || !firstIndex || !secondIndex || !theta || !targetIndices || !discountRatios ||
      !genFlows || !numeraireValues || !discountedFlows)
    return set_error(KOO_ERR_INVALID, "cashFlowDiscounting_partial_gpu: null pointer");
  hipLaunchKernelGGL(k_discounting_partial, dim3((paths + 255) / 256), dim3(256), 0,
                     as_stream(stream), numberFlowsPerPath, genIndices, firstIndex, secondIndex,
                     theta, targetIndices, discountRatios, genFlows, numeraireValues, paths,
                     numberSteps, discountedFlows);
  return check_launch("cashFlowDiscounting_partial_gpu");
}

extern "C" int AggregateFlows_gpu(float* aggregatedFlows, int totalPaths, int numberExerciseDates,
                                  float* genFlows, int batchPaths, int offSetForOutput,
                                  int numberSteps, int* precedingIndex, int start, void* stream) {
  if (batchPaths < 0 || numberSteps < 0 || offSetForOutput < 0 || start < 0 ||
      offSetForOutput + batchPaths > totalPaths)
    return set_error(KOO_ERR_INVALID, "AggregateFlows_gpu: bad sizes");
  if (batchPaths == 0 || numberSteps == 0) return KOO_OK;
  if (!aggregatedFlows || !genFlows || !precedingIndex)
    return set_error(KOO_ERR_INVALID, "AggregateFlows_gpu: null pointer");
  hipLaunchKernelGGL(k_aggregate, dim3((batchPaths + 255) / 256), dim3(256), 0, as_stream(stream),
                     aggregatedFlows, totalPaths, numberExerciseDates, genFlows, batchPaths,
                     offSetForOutput, numberSteps, precedingIndex, start);
  return check_launch("AggregateFlows_gpu");
}

// cashFlowGeneratorEE_gpu<Swap, earlyExerciseNull, LSAMultiExerciseStrategyQuadratic(Cross)_gpu>
// (cashFlowGen_earlyExProduct_gpu.h:323-497; cashFlowGen_earlyEx_product_main.h,
// the instantiations of LSmulti_test.cu:2185-2236 and bermudanIRDPricer.cu:1391-1440):
// the cancellable swap's second pass.  use_cross_terms selects the cross
// strategy; the exercise-value tables (earlyExerciseNull) are unused, as in
// the reference; isExerciseDate is one byte per step (the reference's bool
// array).
extern "C" int cashFlowGeneratorEE_cancellable_swap_gpu(
    int use_cross_terms, float* genFlows1, float* genFlows2, const float* productData,
    const float* /*exerciseValueDataFloat*/, const int* /*exerciseValueDataInt*/,
    const float* exerciseStrategyDataFloat, int exerciseStrategyDataFloat_size,
    const int* exerciseStrategyDataInt, int exerciseStrategyDataInt_size, int paths,
    int numberSteps, int numberExerciseDates, const unsigned char* isExerciseDate,
    const float* rates1, const float* rates2, const float* rates3, const float* strategyVariables,
    int strategyVariables_maxSize, const float* /*forwards*/, const float* /*discRatios*/,
    float* estContinuationValues, int* exerciseIndices, void* stream) {
  const char* who = "cashFlowGeneratorEE_cancellable_swap_gpu";
  if (paths < 0 || numberSteps < 0 || numberExerciseDates < 1 ||
      exerciseStrategyDataInt_size < 3 + numberExerciseDates || exerciseStrategyDataFloat_size < 1 ||
      strategyVariables_maxSize < 1)
    return set_error(KOO_ERR_INVALID, "%s: bad sizes", who);
  if (paths == 0 || numberSteps == 0) return KOO_OK;
  if (!genFlows1 || !genFlows2 || !productData || !exerciseStrategyDataFloat ||
      !exerciseStrategyDataInt || !isExerciseDate || !rates1 || !strategyVariables ||
      !estContinuationValues || !exerciseIndices)
    return set_error(KOO_ERR_INVALID, "%s: null pointer", who);
  hipStream_t st = as_stream(stream);
  const dim3 g((paths + 255) / 256);
  if (use_cross_terms)
    hipLaunchKernelGGL(k_cashflows_ee<true>, g, dim3(256), 0, st, genFlows1, genFlows2,
                       productData, exerciseStrategyDataFloat, exerciseStrategyDataInt, paths,
                       numberSteps, numberExerciseDates, isExerciseDate, rates1, rates2, rates3,
                       strategyVariables, strategyVariables_maxSize, estContinuationValues,
                       exerciseIndices);
  else
    hipLaunchKernelGGL(k_cashflows_ee<false>, g, dim3(256), 0, st, genFlows1, genFlows2,
                       productData, exerciseStrategyDataFloat, exerciseStrategyDataInt, paths,
                       numberSteps, numberExerciseDates, isExerciseDate, rates1, rates2, rates3,
                       strategyVariables, strategyVariables_maxSize, estContinuationValues,
                       exerciseIndices);
  return check_launch(who);
}
