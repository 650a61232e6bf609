// berm_adjoint.hip -- Bermudan payer swaption under a frozen exercise
// strategy: price and pathwise adjoint Greeks, gfx950.
//
// Forward (k_berm_fwd): one lane per path evolves the curve (the shared PC
// step, state in LDS), checkpoints the state every k steps, and at each
// exercise date applies the frozen Longstaff-Schwartz rule (exercise when the
// exercise value is positive, if asked, and exceeds the continuation value of
// the given coefficients).  At exercise it records the step, the deflated
// value and the payoff's adjoint on the end-of-step log-rates,
// d(ev/B)/dl_r = g_r (P_N + K Q_r) / B e^{l_r}, g_r = tau_r / (1 + tau_r f_r),
// Q_r = sum_{m>r} tau_{m-1} P_m (two passes over the alive rates).  A
// wavefront stops once all its paths have exercised.
//
// Backward (k_berm_bwd): the caplet backward's segment recompute and reverse
// step (lmm_step.h), run per wavefront only up to its last exercise step; at a
// path's exercise step the recorded payoff adjoint seeds lambda, and each
// earlier fixing f_s(T_s) inside the numeraire adds -V g_s e^{l_s}.
#include <algorithm>
#include <vector>

#include "berm_device.h"

namespace koo {
namespace {

using namespace lmm;

template <int F>
__global__ void __launch_bounds__(256)
    k_berm_fwd(KOO_MODEL_PARAMS, const BermArgs b, const double* __restrict__ coeffs,
               float* __restrict__ ckpt, float* __restrict__ lamt, int* __restrict__ tau_out,
               float* __restrict__ vout, double* __restrict__ part) {
  KOO_MODEL_LOCAL;
  berm_fwd_block<F>(m, b, CoefGlobal{coeffs}, ckpt, lamt, tau_out, vout, part, blockIdx.x);
}

template <int F>
__global__ void __launch_bounds__(256)
    k_berm_bwd(KOO_MODEL_PARAMS, const float* __restrict__ vols, const float* __restrict__ ckpt,
               const float* __restrict__ lamt, const int* __restrict__ tau_in,
               const float* __restrict__ vin, float* __restrict__ scratch,
               double* __restrict__ part) {
  KOO_MODEL_LOCAL;
  berm_bwd_block<F>(m, vols, ckpt, lamt, tau_in, vin, scratch, part, blockIdx.x);
}

// ---------------------------------------------------------------------------
// Batched portfolio: one launch runs one stage (forward or backward) for
// every product of a portfolio.  The products' descriptors sit in device
// memory; blocks are numbered product after product (the host groups the
// products by rate-count class, one launch per class with LDS sized for the
// class's largest curve, and orders each class by cost, longest first, so
// short products fill the tail), and each block finds its product by a
// binary search over the first-block offsets (uniform per block: scalar
// loads) and dispatches on its factor count.  Every product runs 256-lane
// blocks.
// ---------------------------------------------------------------------------
struct BermItem {
  g_cfloat *z, *f0, *taus, *disp, *A, *vols;
  g_cdouble* coeffs;
  g_float *ckpt, *lamt, *vout, *scratch;
  g_int* tau;
  g_double *part_fwd, *part_bwd, *price_out, *greeks_out;
  float strike;
  int P, N, k, F, block0, blocks;
  BermArgs b;
};

__device__ __forceinline__ const BermItem& find_item(const BermItem* items, int n, int block) {
  return items[find_block_owner(items, n, block, &BermItem::block0)];
}

__device__ __forceinline__ Model item_model(const BermItem& it) {
  return Model{gmem(it.z), cmem(it.f0), cmem(it.taus), cmem(it.disp), cmem(it.A), it.strike,
               it.P, it.N, it.k};
}

__global__ void __launch_bounds__(256)
    k_berm_fwd_batch(const BermItem* __restrict__ items, int n, int block_offset) {
  const int block = blockIdx.x + block_offset;
  const BermItem& it = find_item(items, n, block);
  const Model m = item_model(it);
  const int blk = block - it.block0;
  switch (it.F) {
#define KOO_CASE(F)                                                                            \
  case F:                                                                                      \
    berm_fwd_block<F>(m, it.b, CoefGlobal{gmem(it.coeffs)}, gmem(it.ckpt), gmem(it.lamt),       \
                      gmem(it.tau), gmem(it.vout), gmem(it.part_fwd), blk);                    \
    break;
    KOO_CASE(1) KOO_CASE(2) KOO_CASE(3) KOO_CASE(4) KOO_CASE(5)
#undef KOO_CASE
  }
}

__global__ void __launch_bounds__(256)
    k_berm_bwd_batch(const BermItem* __restrict__ items, int n, int block_offset) {
  const int block = blockIdx.x + block_offset;
  const BermItem& it = find_item(items, n, block);
  const Model m = item_model(it);
  const int blk = block - it.block0;
  switch (it.F) {
#define KOO_CASE(F)                                                                            \
  case F:                                                                                      \
    berm_bwd_block<F>(m, gmem(it.vols), gmem(it.ckpt), gmem(it.lamt), gmem(it.tau),            \
                      gmem(it.vout), gmem(it.scratch), gmem(it.part_bwd), blk);               \
    break;
    KOO_CASE(1) KOO_CASE(2) KOO_CASE(3) KOO_CASE(4) KOO_CASE(5)
#undef KOO_CASE
  }
}

// Per product: price (column 0 of the forward partials) and the 2N Greeks.
// grid (2 * max N + 1, products); block x = 0 is the price.
__global__ void __launch_bounds__(256)
    k_berm_finalize_batch(const BermItem* __restrict__ items, int fwd) {
  const BermItem& it = items[blockIdx.y];
  const double scale = 1.0 / static_cast<double>(it.P);
  if (fwd) {
    if (blockIdx.x == 0)
      finalize_col(gmem(it.part_fwd), it.blocks, 1, 0, scale, gmem(it.price_out));
  } else if (static_cast<int>(blockIdx.x) < 2 * it.N) {
    finalize_col(gmem(it.part_bwd), it.blocks, 2 * it.N, blockIdx.x, scale, gmem(it.greeks_out));
  }
}

}  // namespace
}  // namespace koo

using namespace koo;

extern "C" int bermudan_adjoint_workspace(int paths, int rates, int factors, int checkpoint_every,
                                          size_t* checkpoint_bytes, size_t* scratch_bytes,
                                          size_t* path_bytes, size_t* partial_bytes) {
  if (int e = check_berm("bermudan_adjoint_workspace", paths, rates, factors, checkpoint_every))
    return e;
  const size_t P = static_cast<size_t>(paths);
  const int k = checkpoint_every;
  const size_t nseg = (rates + k - 1) / k;
  const size_t rows = k == 1 ? rates + 1 + 2 * factors : rates + 1;
  // the checkpoints, then the step table the forward copies for the backward
  if (checkpoint_bytes) *checkpoint_bytes = sizeof(float) * (nseg * rows * P + step_table_floats(rates));
  if (scratch_bytes)
    *scratch_bytes = k == 1 ? sizeof(float)
                            : sizeof(float) * static_cast<size_t>(k < rates ? k : rates) *
                                  (rates + 1 + 2 * factors) * P;
  // payoff adjoints (rates x P floats), exercise step (P ints), value (P floats)
  if (path_bytes) *path_bytes = sizeof(float) * (static_cast<size_t>(rates) + 2) * P;
  const size_t bf = (P + fwd_threads(rates) - 1) / fwd_threads(rates);
  const size_t bb = (P + bwd_threads(rates) - 1) / bwd_threads(rates);
  const size_t need = bf > bb * 2 * rates ? bf : bb * 2 * rates;
  if (partial_bytes) *partial_bytes = sizeof(double) * need;
  return KOO_OK;
}

extern "C" int bermudan_frozen_fwd(int paths, int rates, int factors, int checkpoint_every,
                                   const float* variates, const float* initial_rates,
                                   const float* taus, const float* displacements,
                                   const float* pseudo_roots, float strike, int n_ex,
                                   const int* exercise_steps, const double* coefficients,
                                   const double* var_shift, const double* var_scale, int itm_only,
                                   float* checkpoints, float* path_buffers, double* partials,
                                   double* price_out, void* stream) {
  if (int e = check_berm("bermudan_frozen_fwd", paths, rates, factors, checkpoint_every)) return e;
  if (!variates || !initial_rates || !taus || !displacements || !pseudo_roots || !coefficients ||
      !checkpoints || !path_buffers || !partials || !price_out)
    return set_error(KOO_ERR_INVALID, "bermudan_frozen_fwd: null pointer");
  BermArgs b;
  if (int e = make_berm_args("bermudan_frozen_fwd", rates, n_ex, exercise_steps, var_shift,
                             var_scale, itm_only, &b))
    return e;
  const size_t P = static_cast<size_t>(paths);
  float* lamt = path_buffers;
  int* tau = reinterpret_cast<int*>(path_buffers + static_cast<size_t>(rates) * P);
  float* vout = path_buffers + static_cast<size_t>(rates + 1) * P;
  const int threads = fwd_threads(rates);
  const int blocks = (paths + threads - 1) / threads;
  const size_t lds = fwd_lds(rates, threads);
  hipStream_t st = as_stream(stream);
  switch (factors) {
#define KOO_CASE(F)                                                                            \
  case F:                                                                                      \
    hipLaunchKernelGGL(k_berm_fwd<F>, dim3(blocks), dim3(threads), lds, st, variates,          \
                       initial_rates, taus, displacements, pseudo_roots, strike, paths, rates, \
                       checkpoint_every, b, coefficients, checkpoints, lamt, tau, vout,        \
                       partials);                                                              \
    break;
    KOO_CASE(1) KOO_CASE(2) KOO_CASE(3) KOO_CASE(4) KOO_CASE(5) KOO_CASE(6) KOO_CASE(7)
    KOO_CASE(8)
#undef KOO_CASE
  }
  hipLaunchKernelGGL(k_finalize_cols, dim3(1), dim3(256), 0, st, partials, blocks, 1,
                     1.0 / static_cast<double>(paths), price_out);
  return check_launch("bermudan_frozen_fwd");
}

extern "C" int bermudan_frozen_bwd(int paths, int rates, int factors, int checkpoint_every,
                                   const float* variates, const float* initial_rates,
                                   const float* taus, const float* displacements,
                                   const float* pseudo_roots, const float* vols, float strike,
                                   const float* checkpoints, const float* path_buffers,
                                   float* segment_scratch, double* partials, double* greeks_out,
                                   void* stream) {
  if (int e = check_berm("bermudan_frozen_bwd", paths, rates, factors, checkpoint_every)) return e;
  if (!variates || !initial_rates || !taus || !displacements || !pseudo_roots || !vols ||
      !checkpoints || !path_buffers || !segment_scratch || !partials || !greeks_out)
    return set_error(KOO_ERR_INVALID, "bermudan_frozen_bwd: null pointer");
  const size_t P = static_cast<size_t>(paths);
  const float* lamt = path_buffers;
  const int* tau = reinterpret_cast<const int*>(path_buffers + static_cast<size_t>(rates) * P);
  const float* vin = path_buffers + static_cast<size_t>(rates + 1) * P;
  const int threads = bwd_threads(rates);
  const int blocks = (paths + threads - 1) / threads;
  const size_t lds = bwd_lds(rates, threads, false);
  hipStream_t st = as_stream(stream);
  switch (factors) {
#define KOO_CASE(F)                                                                            \
  case F:                                                                                      \
    hipLaunchKernelGGL(k_berm_bwd<F>, dim3(blocks), dim3(threads), lds, st, variates,          \
                       initial_rates, taus, displacements, pseudo_roots, strike, paths, rates, \
                       checkpoint_every, vols, checkpoints, lamt, tau, vin, segment_scratch,   \
                       partials);                                                              \
    break;
    KOO_CASE(1) KOO_CASE(2) KOO_CASE(3) KOO_CASE(4) KOO_CASE(5) KOO_CASE(6) KOO_CASE(7)
    KOO_CASE(8)
#undef KOO_CASE
  }
  hipLaunchKernelGGL(k_finalize_cols, dim3(2 * rates), dim3(256), 0, st, partials, blocks,
                     2 * rates, 1.0 / static_cast<double>(paths), greeks_out);
  return check_launch("bermudan_frozen_bwd");
}

// ---------------------------------------------------------------------------
// Batched portfolio entry points.
// ---------------------------------------------------------------------------
extern "C" size_t bermudan_portfolio_descriptor_bytes(int n_products) {
  return n_products > 0 ? sizeof(BermItem) * static_cast<size_t>(n_products) : 0;
}

extern "C" int bermudan_portfolio_prepare(int n_products, const koo_berm_product* products,
                                          void* descriptors_device, koo_portfolio_plan* plan) {
  const char* who = "bermudan_portfolio_prepare";
  if (n_products < 1 || !products || !descriptors_device || !plan)
    return set_error(KOO_ERR_INVALID, "%s: bad arguments", who);
  std::vector<BermItem> items(n_products);
  std::vector<double> cost(n_products);
  int max_n = 0;
  for (int i = 0; i < n_products; ++i) {
    const koo_berm_product& q = products[i];
    if (int e = check_berm(who, q.paths, q.rates, q.factors, q.checkpoint_every)) return e;
    if (q.factors > 5 || q.rates > 45)
      return set_error(KOO_ERR_INVALID,
                       "%s: product %d: batched kernels take factors <= 5 and rates <= 45", who,
                       i);
    if (!q.variates || !q.initial_rates || !q.taus || !q.displacements || !q.pseudo_roots ||
        !q.vols || !q.coefficients || !q.checkpoints || !q.path_buffers ||
        !q.segment_scratch || !q.partials || !q.price_out || !q.greeks_out)
      return set_error(KOO_ERR_INVALID, "%s: product %d: null pointer", who, i);
    BermItem& it = items[i];
    it = BermItem{};
    if (int e = make_berm_args(who, q.rates, q.n_ex, q.exercise_steps, q.var_shift, q.var_scale,
                               q.itm_only, &it.b))
      return e;
    const size_t P = static_cast<size_t>(q.paths);
    it.z = (g_cfloat*)q.variates;
    it.f0 = (g_cfloat*)q.initial_rates;
    it.taus = (g_cfloat*)q.taus;
    it.disp = (g_cfloat*)q.displacements;
    it.A = (g_cfloat*)q.pseudo_roots;
    it.vols = (g_cfloat*)q.vols;
    it.coeffs = (g_cdouble*)q.coefficients;
    it.ckpt = (g_float*)q.checkpoints;
    it.lamt = (g_float*)q.path_buffers;
    it.tau = (g_int*)(q.path_buffers + static_cast<size_t>(q.rates) * P);
    it.vout = (g_float*)(q.path_buffers + static_cast<size_t>(q.rates + 1) * P);
    it.scratch = (g_float*)q.segment_scratch;
    it.part_fwd = it.part_bwd = (g_double*)q.partials;
    it.price_out = (g_double*)q.price_out;
    it.greeks_out = (g_double*)q.greeks_out;
    it.strike = q.strike;
    it.P = q.paths;
    it.N = q.rates;
    it.k = q.checkpoint_every;
    it.F = q.factors;
    it.blocks = (q.paths + 255) / 256;
    // work ~ paths x rate updates up to the last exercise step
    double u = 0.0;
    for (int s = 0; s <= it.b.last; ++s) u += q.rates - s;
    cost[i] = u * q.paths * (q.factors + 4);
    if (q.rates > max_n) max_n = q.rates;
  }
  std::vector<int> order(n_products);
  for (int i = 0; i < n_products; ++i) order[i] = i;
  // rate class first (one launch per class), then most expensive first
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    const int ca = rate_class(items[a].N), cb = rate_class(items[b].N);
    return ca != cb ? ca < cb : cost[a] > cost[b];
  });
  std::vector<BermItem> sorted(n_products);
  int64_t block0 = 0;
  *plan = koo_portfolio_plan{};
  for (int i = 0; i < n_products; ++i) {
    sorted[i] = items[order[i]];
    sorted[i].block0 = static_cast<int>(block0);
    const int c = rate_class(sorted[i].N);
    if (plan->classes == 0 || rate_class(sorted[i - 1].N) != c) {
      plan->class_first_block[plan->classes] = static_cast<int>(block0);
      ++plan->classes;
    }
    plan->class_blocks[plan->classes - 1] += sorted[i].blocks;
    if (sorted[i].N > plan->class_max_rates[plan->classes - 1])
      plan->class_max_rates[plan->classes - 1] = sorted[i].N;
    block0 += sorted[i].blocks;
  }
  if (block0 > 0x7fffffff) return set_error(KOO_ERR_INVALID, "%s: too many blocks", who);
  if (hipMemcpy(descriptors_device, sorted.data(), sizeof(BermItem) * n_products,
                hipMemcpyHostToDevice) != hipSuccess)
    return set_error(KOO_ERR_HIP, "%s: descriptor copy failed", who);
  plan->products = n_products;
  plan->blocks = static_cast<int>(block0);
  plan->max_rates = max_n;
  return KOO_OK;
}

static int portfolio_launch(const char* who, bool fwd, const koo_portfolio_plan* plan,
                            const void* descriptors_device, void* stream) {
  if (!plan || !descriptors_device || plan->products < 1 || plan->blocks < 1 ||
      plan->max_rates < 1 || plan->max_rates > 45 || plan->classes < 1 || plan->classes > 4)
    return set_error(KOO_ERR_INVALID, "%s: bad plan", who);
  const BermItem* items = static_cast<const BermItem*>(descriptors_device);
  const int N = plan->max_rates;
  hipStream_t st = as_stream(stream);
  for (int c = 0; c < plan->classes; ++c) {
    const int nc = plan->class_max_rates[c], b0 = plan->class_first_block[c];
    const int nb = plan->class_blocks[c];
    if (nc < 1 || nc > N || nb < 1 || b0 < 0 || b0 + nb > plan->blocks)
      return set_error(KOO_ERR_INVALID, "%s: bad plan class %d", who, c);
    if (fwd)
      hipLaunchKernelGGL(k_berm_fwd_batch, dim3(nb), dim3(256), fwd_lds(nc, 256), st, items,
                         plan->products, b0);
    else
      hipLaunchKernelGGL(k_berm_bwd_batch, dim3(nb), dim3(256), bwd_lds(nc, 256, false), st, items,
                         plan->products, b0);
  }
  hipLaunchKernelGGL(k_berm_finalize_batch, dim3(fwd ? 1 : 2 * N, plan->products), dim3(256), 0,
                     st, items, fwd ? 1 : 0);
  return check_launch(who);
}

extern "C" int bermudan_portfolio_fwd(const koo_portfolio_plan* plan,
                                      const void* descriptors_device, void* stream) {
  return portfolio_launch("bermudan_portfolio_fwd", true, plan, descriptors_device, stream);
}

extern "C" int bermudan_portfolio_bwd(const koo_portfolio_plan* plan,
                                      const void* descriptors_device, void* stream) {
  return portfolio_launch("bermudan_portfolio_bwd", false, plan, descriptors_device, stream);
}
