// ls.hip -- Longstaff-Schwartz regression on gfx950.
//
// * k_berm_paths: one lane per path evolves the LMM (the shared
//   predictor-corrector step) up to the last exercise date and writes, per
//   exercise date, the basis variables (co-terminal swap rate, first alive
//   forward), the Bermudan payer swaption exercise value and the spot
//   numeraire.  The rate cube is never written.
// * k_gram: sum over paths of the outer product of up to 16 path-minor rows
//   (X^T X and X^T y in one Gram matrix), on MFMA v_mfma_f32_16x16x4f32 with
//   FP32 accumulation inside a block and FP64 block partials.
// * LS first pass, in rounds (one per exercise date, last date first, then
//   date 0's decision and the price): every block sums the previous round's
//   FP64 block partials and solves that date's normal equations itself (same
//   inputs and order, so identical coefficients), then applies the later
//   date's exercise decision and numeraire deflation to its paths' running
//   values and accumulates this date's normal equations (optionally over
//   in-the-money paths only) on MFMA into its block partials.  One launch per
//   round (k_ls_round_single, the default) or one persistent launch
//   (k_ls_first_pass; the partials exchanged through words armed with a
//   sentinel, no counters).  The portfolio's passes run the same per-block
//   code (k_ls_portfolio, k_ls_round).
// * k_ls_second: the frozen-strategy pass (price under given coefficients).
// * the reference's small LS kernels (basis variables, exercise values, basis
//   expansion, continuation values, update, deflation).
#include <algorithm>
#include <cstdlib>
#include <vector>

#include "ls_device.h"

namespace koo {
namespace {

using namespace lmm;

// ---------------------------------------------------------------------------
// Bermudan payer swaption path quantities.
// ---------------------------------------------------------------------------
template <int F>
__global__ void __launch_bounds__(256)
    k_berm_paths(KOO_MODEL_PARAMS, const ExMap ex, float* __restrict__ bv, float* __restrict__ evo,
                 float* __restrict__ numo) {
  KOO_MODEL_LOCAL;
  berm_paths_block<F>(m, ex, bv, evo, numo, blockIdx.x);
}

// ---------------------------------------------------------------------------
// Gram matrix of up to 16 path-minor rows (optionally masked), FP64 partials
// per block into part[block][n*n].
// ---------------------------------------------------------------------------
struct RowSet {
  const float* row[kMaxGram];
};

template <int GROUPS>
__global__ void __launch_bounds__(kGramThreads)
    k_gram(const RowSet rows, int n, const float* __restrict__ mask, int P,
           double* __restrict__ part) {
  __shared__ float lds_d[(kGramThreads / 64) * 256];
  constexpr int kPathsPerOp = 4 * GROUPS;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const int j = GROUPS == 2 ? (lane & 7) : (lane & 15);
  const int sub = GROUPS == 2 ? ((lane >> 3) & 1) * 4 + (lane >> 4) : (lane >> 4);
  const float* __restrict__ src = rows.row[j < n ? j : 0];
  floatx4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
  const int stride = gridDim.x * nw * 64;
  for (int base = (blockIdx.x * nw + wave) * 64; base < P; base += stride) {
#pragma unroll 4
    for (int q = 0; q < 64 / kPathsPerOp; ++q) {
      const int p = base + q * kPathsPerOp + sub;
      float a = 0.0f;
      if (j < n && p < P) {
        a = src[p];
        if (mask && mask[p] == 0.0f) a = 0.0f;
      }
      acc = mfma16x16x4(a, a, acc);
    }
  }
  block_gram<GROUPS>(acc, n, lds_d, part, blockIdx.x, gridDim.x);
}

// Sums the partials of k_gram and writes the n x n result (FP64 and, when
// asked, FP32 scaled by alpha plus beta times the old FP32 value) for the
// (r0, c0, nr, nc) block of it, row-major with leading dimension ld.
__global__ void __launch_bounds__(kGramThreads)
    k_gram_finish(const double* __restrict__ part, int nb, int n, int r0, int c0, int nr, int nc,
                  int ld, double* __restrict__ out64, float* __restrict__ out32, float alpha,
                  float beta, int transpose) {
  __shared__ double sum[kMaxGram * kMaxGram];
  reduce_partials(part, nb, n * n, sum);
  for (int t = threadIdx.x; t < nr * nc; t += blockDim.x) {
    const int i = t / nc, j = t - i * nc;
    const double v = sum[(r0 + i) * n + c0 + j];
    const int o = transpose ? j * ld + i : i * ld + j;
    if (out64) out64[o] = v;
    if (out32) out32[o] = alpha * static_cast<float>(v) + (beta != 0.0f ? beta * out32[o] : 0.0f);
  }
}

// ---------------------------------------------------------------------------
// Frozen-strategy pass (SecondPassPriceLSAMultiUsingAggregatedFlows_gold,
// gold/MultiLS_regression.h:237-289 at depth 0): walk the dates, add the
// deflated cash flow until the strategy exercises, then the deflated exercise
// value.  The per-date variables and shifts travel as a by-value kernel
// argument (about 3 KB, within the 4 KB argument limit).
struct AllDates {
  int E, P, max_vars, maxd, cross, itm_only;
  signed char v[kMaxRates], k[kMaxRates];
  float sh[kMaxRates][kMaxBasisVars], sc[kMaxRates][kMaxBasisVars];
};

// The arguments of round r (date e = E - 1 - r; e = -1 for round E).
__device__ __forceinline__ DateArgs round_args(const AllDates& d, int r) {
  DateArgs a{};
  const int e = d.E - 1 - r;
  a.e = e;
  a.E = d.E;
  a.P = d.P;
  a.max_vars = d.max_vars;
  a.maxd = d.maxd;
  a.cross = d.cross;
  a.itm_only = d.itm_only;
  if (e >= 0) {
    a.v = d.v[e];
    a.k = d.k[e];
    for (int i = 0; i < kMaxBasisVars; ++i) {
      a.sh[i] = d.sh[e][i];
      a.sc[i] = d.sc[e][i];
    }
  }
  if (e + 1 < d.E) {
    a.vn = d.v[e + 1];
    a.kn = d.k[e + 1];
    for (int i = 0; i < kMaxBasisVars; ++i) {
      a.shn[i] = d.sh[e + 1][i];
      a.scn[i] = d.sc[e + 1][i];
    }
  }
  return a;
}

// Entries of round r's partials: the Gram triangle of its date, or the price.
__device__ __forceinline__ int round_entries(const AllDates& d, int r) {
  const int e = d.E - 1 - r;
  return e >= 0 ? tri_count(d.k[e] + 1) : 1;
}

// One block's round r of the first pass (launch per round): the later
// date's coefficients from the previous round's partials (every block solves;
// block 0 stores them), then the path work.
template <int GROUPS>
__global__ void __launch_bounds__(kGramThreads)
    k_ls_round_single(const AllDates dv, int r, const float* __restrict__ bv,
                      const float* __restrict__ cash, const float* __restrict__ ev,
                      const float* __restrict__ num, double* __restrict__ coeffs,
                      float* __restrict__ Y, double* __restrict__ xbuf) {
  __shared__ double tri[kTriMax];
  __shared__ double lds_d[(kGramThreads / 64) * 256];
  __shared__ double gram[kMaxGram * kMaxGram];
  __shared__ double scratch[kSolveScratch];
  __shared__ double beta[kMaxGram];
  const AllDates& d = dv;
  const int nb = gridDim.x;
  const DateArgs a = round_args(d, r);
  float unused = 0.0f;
  const PathIn first =
      load_path(a, bv, cash, ev, num, Y, min(static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x), d.P - 1));
  if (r > 0) {
    sum_tri<false>(xbuf + ls_slot(r - 1, nb), nb, round_entries(d, r - 1), tri, nullptr);
    solve_tri(tri, a.kn, gram, scratch, beta);
    if (blockIdx.x == 0 && threadIdx.x < d.maxd)
      coeffs[static_cast<size_t>(a.e + 1) * d.maxd + threadIdx.x] =
          threadIdx.x < a.kn ? beta[threadIdx.x] : 0.0;
  }
  if (a.e >= 0)
    ls_date_work<GROUPS, false, kMaxGram>(a, bv, cash, ev, num, beta, Y, false, unused,
                                          xbuf + ls_slot(r, nb), blockIdx.x, nb, first, lds_d);
  else
    ls_final_work<false, kMaxGram>(a, bv, cash, ev, num, beta, Y, false, 0.0f,
                                   xbuf + ls_slot(r, nb), blockIdx.x, nb, first);
}

// The whole first pass in one persistent launch (at most as many blocks as
// can be resident, every basis of at most 7 functions).  The round's
// partials are exchanged through agent-coherent words armed with kUnsetBits
// (k_fill_u64 before the launch): a block stores its column, then reads every
// block's column of the round until none is unset.  A block re-arms its own
// words of round r - 2 once it has read round r - 1 (every block stored
// round r - 1 after reading round r - 2), and waits for those stores before
// storing round r, so a reader of round r + 1 never sees round r - 2's value.
// The next round's first chunk of path data is loaded before the wait.  If a
// wait gives up: NaN price and sync[0] = 1.
__global__ void __launch_bounds__(kGramThreads)
    k_ls_first_pass(const AllDates dv, const float* __restrict__ bv,
                    const float* __restrict__ cash, const float* __restrict__ ev,
                    const float* __restrict__ num, double* __restrict__ coeffs,
                    float* __restrict__ Y, double* __restrict__ xbuf,
                    unsigned* __restrict__ sync, double* __restrict__ price_out) {
  __shared__ int flag;
  __shared__ double tri[kTriMax];
  __shared__ double lds_d[(kGramThreads / 64) * 256];
  __shared__ double gram[kMaxGram * kMaxGram];
  __shared__ double scratch[kSolveScratch];
  __shared__ double beta[kMaxGram];
  const AllDates& d = dv;
  const int nb = gridDim.x;
  const int p0 = min(static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x), d.P - 1);
  // one chunk of paths per lane: the running value stays in a register
  float yr = 0.0f;
  const bool yin = static_cast<long long>(nb) * blockDim.x >= d.P;
  const float* Yl = yin ? nullptr : Y;
  PathIn first = load_path(round_args(d, 0), bv, cash, ev, num, Yl, p0);
  for (int r = 0; r <= d.E; ++r) {
    const DateArgs a = round_args(d, r);
    if (r > 0) {
      if (!sum_tri<true>(xbuf + ls_slot(r - 1, nb), nb, round_entries(d, r - 1), tri, &flag))
        goto fail;
      if (r >= 2) rearm_tri(xbuf + ls_slot(r - 2, nb), nb, round_entries(d, r - 2), blockIdx.x);
      solve_tri(tri, a.kn, gram, scratch, beta);
      // the re-arming stores performed before any wave stores this round's
      // partials (after the barriers in the work)
      if (r >= 2) __builtin_amdgcn_s_waitcnt(0x0F70);
      if (blockIdx.x == 0 && threadIdx.x < d.maxd)
        coeffs[static_cast<size_t>(a.e + 1) * d.maxd + threadIdx.x] =
            threadIdx.x < a.kn ? beta[threadIdx.x] : 0.0;
    }
    double* out = xbuf + ls_slot(r, nb);
    if (a.e >= 0) {
      ls_date_work<2, true, 8>(a, bv, cash, ev, num, beta, Y, yin, yr, out, blockIdx.x, nb, first,
                               lds_d);
      first = load_path(round_args(d, r + 1), bv, cash, ev, num, Yl, p0);
    } else {
      ls_final_work<true, 8>(a, bv, cash, ev, num, beta, Y, yin, yr, out, blockIdx.x, nb, first);
      if (blockIdx.x == 0) {
        if (!sum_tri<true>(out, nb, 1, tri, &flag)) goto fail;
        finalize_col<true>(out, nb, 1, 0, 1.0 / static_cast<double>(d.P), price_out);
      }
    }
  }
  return;
fail:
  if (threadIdx.x == 0) {
    sync[0] = 1u;
    *price_out = __builtin_nan("");
  }
}

// Frozen-strategy pass: each path walks the dates until it exercises.  The
// walk stops at a data-dependent date, so loading a date's inputs only after
// the previous decision would make a chain of dependent memory round trips;
// the inputs of kSecondChunk dates are loaded together ahead of their
// decisions instead (a path that stops early leaves the rest unused).  Per
// path the same operations in the same order as a date-by-date walk.
constexpr int kSecondChunk = 4;

__global__ void __launch_bounds__(kLsThreads)
    k_ls_second(const AllDates dv, const float* __restrict__ bv,
                const float* __restrict__ cash, const float* __restrict__ ev,
                const float* __restrict__ num, const double* __restrict__ coeffs,
                double* __restrict__ part) {
  __shared__ double red[4];
  const AllDates* d = &dv;
  const size_t P = static_cast<size_t>(d->P);
  double s = 0.0;
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < d->P; p += gridDim.x * blockDim.x) {
    double acc = 0.0;
    bool done = false;
    for (int e0 = 0; e0 < d->E && !done; e0 += kSecondChunk) {
      float xr[kSecondChunk][kMaxBasisVars], exv[kSecondChunk], nm[kSecondChunk],
          cf[kSecondChunk];
#pragma unroll
      for (int c = 0; c < kSecondChunk; ++c) {
        const int e = e0 + c;
        const bool in = e < d->E;
        const size_t o = static_cast<size_t>(in ? e : 0) * P + p;
        const int v = in ? d->v[e] : 0;
#pragma unroll
        for (int i = 0; i < kMaxBasisVars; ++i)
          xr[c][i] = i < v ? bv[(static_cast<size_t>(e) * d->max_vars + i) * P + p] : 0.0f;
        exv[c] = in ? ev[o] : 0.0f;
        nm[c] = in ? num[o] : 1.0f;
        cf[c] = (in && cash) ? cash[o] : 0.0f;
      }
#pragma unroll
      for (int c = 0; c < kSecondChunk; ++c) {
        const int e = e0 + c;
        if (done || e >= d->E) break;
        const float cont = continuation_raw(xr[c], d->v[e], d->k[e], d->cross != 0, d->sh[e],
                                            d->sc[e], coeffs + static_cast<size_t>(e) * d->maxd);
        const bool allowed = !d->itm_only || exv[c] > 0.0f;
        if (allowed && cont - exv[c] < 0.0f) {
          acc += static_cast<double>(exv[c]) / static_cast<double>(nm[c]);
          done = true;
        } else if (cash) {
          acc += static_cast<double>(cf[c]) / static_cast<double>(nm[c]);
        }
      }
    }
    s += acc;
  }
  block_sum_to(s, red, part + blockIdx.x);
}

// ---------------------------------------------------------------------------
// The reference's small LS kernels.
// ---------------------------------------------------------------------------

// basis_function_kernel with basisVariableExample / basisVariableLog_gpu
// (LS_Basis_gpu.h:24-94, LS_basis_examples_gpu.h:28-147): three variables per
// exercise date, (E, 3, totalPaths).
__global__ void k_basis_variables(int use_log, const float* __restrict__ shifts,
                                  const float* __restrict__ r1, const float* __restrict__ r2,
                                  const float* __restrict__ r3, int batch, int total, int offset,
                                  int n_ex, float* __restrict__ out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= batch) return;
  float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f;
  if (use_log) {
    s0 = shifts[0];
    s1 = shifts[1];
    s2 = shifts[2];
  }
  for (int e = 0; e < n_ex; ++e) {
    const size_t in = static_cast<size_t>(e) * batch + p;
    float a = r1[in], b = r2[in], c = r3[in];
    if (use_log) {
      a = logf(a + s0);
      b = logf(b + s1);
      c = logf(c + s2);
    }
    float* o = out + static_cast<size_t>(e) * 3 * total + offset + p;
    o[0] = a;
    o[static_cast<size_t>(total)] = b;
    o[2 * static_cast<size_t>(total)] = c;
  }
}

// exercise_value_kernel with earlyExerciseBermudanPayerSwaption
// (early_exercise_value_generator_gpu.h:38-95, exercise_values_examples.h:49-78).
__global__ void k_exercise_values_bermudan(const float* __restrict__ aux, const float* __restrict__ r1,
                                           const float* __restrict__ r2, int batch, int total,
                                           int offset, int n_ex, float* __restrict__ out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= batch) return;
  const float K = aux[0];
  for (int e = 0; e < n_ex; ++e) {
    const size_t in = static_cast<size_t>(e) * batch + p;
    const float v = (r1[in] - K) * r2[in];
    out[static_cast<size_t>(e) * total + offset + p] = v > 0.0f ? v : 0.0f;
  }
}

// basis_function_expansion_kernel (LS_gpu.h:232-272): (v, P) -> (k, P).
__global__ void k_expand_basis(const float* __restrict__ in, int P, int v, int cross,
                               float* __restrict__ out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  const size_t sP = static_cast<size_t>(P);
  float x[kMaxBasisVars] = {0.0f, 0.0f, 0.0f, 0.0f};
  for (int i = 0; i < v; ++i) x[i] = in[i * sP + p];
  const int k = basis_count(v, cross != 0);
  for (int j = 0; j < k; ++j)
    out[j * sP + p] = basis_value(j, v, cross != 0, x[0], x[1], x[2], x[3]);
}

// continuation_net_value_evaluation_kernel (LS_gpu.cu:56-85).
__global__ void k_net_continuation(int P, int k, const float* __restrict__ w,
                                   const float* __restrict__ bf, const float* __restrict__ ev,
                                   float shift, float* __restrict__ out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  const size_t sP = static_cast<size_t>(P);
  float c = shift;
  for (int i = 0; i < k; ++i) c += w[i] * bf[i * sP + p];
  out[p] = c - ev[p];
}

// exercise_updating_LS_multi_kernel (LS_gpu.cu:312-400): nested regressions
// with cut bands, exercise when the net continuation value is negative.
__global__ void k_update_multi(int P, int k, int row, int depth_max, const float* __restrict__ w,
                               const float* __restrict__ lo, const float* __restrict__ hi,
                               const float* __restrict__ bf, float* __restrict__ cont_values,
                               const float* __restrict__ ev, float shift,
                               float* __restrict__ est, int* __restrict__ exercises) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  const size_t sP = static_cast<size_t>(P);
  const float exv = ev[p];
  float c = 0.0f;
  bool exercise = false;
  for (int depth = 0;; ++depth) {
    const float* wd = w + static_cast<size_t>(depth) * row;
    c = shift;
    for (int i = 0; i < k; ++i) c += wd[i] * bf[i * sP + p];
    const float net = c - exv;
    if (depth == depth_max - 1 || net < lo[depth] || net > hi[depth]) {
      exercise = net < 0.0f;
      break;
    }
  }
  if (exercise) cont_values[p] = exv;
  exercises[p] = exercise ? 1 : -1;
  est[p] = c;
}

// numeraire_deflating_kernel (LS_gpu.cu:246-262).
__global__ void k_deflate(float* __restrict__ y, const float* __restrict__ num_new,
                          const float* __restrict__ num_old, int P) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  y[p] *= num_new[p] / num_old[p];
}

// ---------------------------------------------------------------------------
// Host helpers.
// ---------------------------------------------------------------------------
int berm_threads(int N) { return N <= 50 ? 256 : (N <= 75 ? 128 : 64); }

int ls_blocks(int paths) {
  const int want = (paths + kLsThreads - 1) / kLsThreads;
  const int cap = 2 * device_cu_count();
  return want < 1 ? 1 : (want > cap ? cap : want);
}

// Blocks of the 1024-lane MFMA kernels: one per CU at most, so the
// per-date partials stay at most CUs x 256 doubles.
int gram_blocks(int paths) {
  const int want = (paths + kGramThreads - 1) / kGramThreads;
  const int cap = device_cu_count();
  return want < 1 ? 1 : (want > cap ? cap : want);
}

size_t align256(size_t x) { return (x + 255) & ~static_cast<size_t>(255); }

// Blocks of the first pass: gram_blocks(paths), capped at what can be
// resident at once when the persistent kernel can run at all (so both
// launch forms use the same blocks); with fewer than one resident block per
// CU the per-round launches use gram_blocks(paths).
int ls_pass_blocks(int paths, bool* persistent_ok) {
  int per_cu = 0;
  const int want = gram_blocks(paths);
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_ls_first_pass, kGramThreads, 0) !=
          hipSuccess ||
      per_cu < 1) {
    *persistent_ok = false;
    return want;
  }
  *persistent_ok = true;
  const int cap = per_cu * device_cu_count();
  return want < cap ? want : cap;
}

// Workspace: running continuation values (P floats); block partials of the
// reference-named Gram kernels and the second pass (kMaxGram^2 doubles per
// block); the first pass's three partial slots (kTriMax doubles per block
// each); the persistent first pass's error word.
struct LsLayout {
  size_t y, part, xbuf, bar, total;
};

LsLayout ls_layout(int paths) {
  LsLayout l;
  l.y = 0;
  l.part = align256(sizeof(float) * static_cast<size_t>(paths));
  const int nb = ls_blocks(paths) > gram_blocks(paths) ? ls_blocks(paths) : gram_blocks(paths);
  l.xbuf = l.part + align256(sizeof(double) * kMaxGram * kMaxGram * nb);
  l.bar = l.xbuf + align256(sizeof(double) * 3 * kTriMax * gram_blocks(paths));
  l.total = l.bar + 256;
  return l;
}

int check_ls(const char* who, int n_ex, const int* ves, int max_vars, int paths, int use_cross) {
  if (n_ex < 1 || n_ex > kMaxRates || paths < 1 || max_vars < 0 || max_vars > kMaxBasisVars ||
      !ves)
    return set_error(KOO_ERR_INVALID,
                     "%s: need 1<=exercise dates<=%d, paths>=1, 0<=max_vars<=%d (got %d, %d, %d)",
                     who, kMaxRates, kMaxBasisVars, n_ex, paths, max_vars);
  for (int e = 0; e < n_ex; ++e)
    if (ves[e] < 0 || ves[e] > max_vars)
      return set_error(KOO_ERR_INVALID, "%s: vars_each_step[%d] = %d outside [0, %d]", who, e,
                       ves[e], max_vars);
  if (basis_count(max_vars, use_cross != 0) + 1 > kMaxGram)
    return set_error(KOO_ERR_INVALID, "%s: %d basis functions exceed %d", who,
                     basis_count(max_vars, use_cross != 0), kMaxGram - 1);
  return KOO_OK;
}

void fill_norm(const float* shift, const float* scale, int e, int v, int max_vars, float* sh,
               float* sc) {
  for (int i = 0; i < kMaxBasisVars; ++i) {
    sh[i] = (i < v && shift) ? shift[e * max_vars + i] : 0.0f;
    sc[i] = (i < v && scale) ? scale[e * max_vars + i] : 1.0f;
  }
}

// Gram matrix of `n` rows into the workspace partials, then the requested
// block of it; returns the status.
int gram_launch(const char* who, const float* const* rows, int n, const float* mask, int paths,
                double* part, int nb, int r0, int c0, int nr, int nc, int ld, double* out64,
                float* out32, float alpha, float beta, int transpose, hipStream_t st) {
  RowSet rs{};
  for (int i = 0; i < n; ++i) rs.row[i] = rows[i];
  for (int i = n; i < kMaxGram; ++i) rs.row[i] = rows[0];
  if (n <= 8)
    hipLaunchKernelGGL(k_gram<2>, dim3(nb), dim3(kGramThreads), 0, st, rs, n, mask, paths, part);
  else
    hipLaunchKernelGGL(k_gram<1>, dim3(nb), dim3(kGramThreads), 0, st, rs, n, mask, paths, part);
  hipLaunchKernelGGL(k_gram_finish, dim3(1), dim3(kGramThreads), 0, st, part, nb, n, r0, c0, nr, nc,
                     ld, out64, out32, alpha, beta, transpose);
  return check_launch(who);
}

// ---------------------------------------------------------------------------
// Batched portfolio: Bermudan path quantities of every product in one
// launch, then one launch per date rank for the LS first passes (round r
// runs date E - 1 - r of each product, or its date-0 decision when r == E).
// Blocks are numbered product after product, in both kernels; each block
// finds its product by binary search over the first-block offsets.
// ---------------------------------------------------------------------------
constexpr int kLsBatchMaxDates = 40;

struct LsItem {
  g_cfloat *z, *f0, *taus, *disp, *A;
  g_float *bv, *ev, *num, *Y;
  g_double *xbuf, *coeffs, *price_out;  // xbuf: the LS partial slots
  float strike;
  int P, N, F, E, itm_only;
  int paths_block0, paths_blocks, ls_block0, ls_blocks;
  ExMap ex;
  float sh[kLsBatchMaxDates][2], sc[kLsBatchMaxDates][2];
};

__global__ void __launch_bounds__(256)
    k_berm_paths_batch(const LsItem* __restrict__ items, int n, int block_offset) {
  const int block = blockIdx.x + block_offset;
  const LsItem& it = items[find_block_owner(items, n, block, &LsItem::paths_block0)];
  const Model m{gmem(it.z), cmem(it.f0), cmem(it.taus), cmem(it.disp), cmem(it.A), it.strike,
                it.P, it.N, 1};
  const int blk = block - it.paths_block0;
  switch (it.F) {
#define KOO_CASE(F)                                                                            \
  case F:                                                                                      \
    berm_paths_block<F>(m, it.ex, gmem(it.bv), gmem(it.ev), gmem(it.num), blk);                \
    break;
    KOO_CASE(1) KOO_CASE(2) KOO_CASE(3) KOO_CASE(4) KOO_CASE(5)
#undef KOO_CASE
  }
}

// The arguments of a product's round r (two variables per date, the cross
// basis: 6 functions; date e = E - 1 - r, e = -1 for round E).
__device__ __forceinline__ DateArgs item_round_args(const LsItem& it, int r) {
  DateArgs a{};
  const int e = it.E - 1 - r;
  a.e = e;
  a.E = it.E;
  a.P = it.P;
  a.max_vars = 2;
  a.maxd = 6;
  a.cross = 1;
  a.itm_only = it.itm_only;
  if (e >= 0) {
    a.v = 2;
    a.k = 6;
    for (int i = 0; i < kMaxBasisVars; ++i) {
      a.sh[i] = i < 2 ? it.sh[e][i] : 0.0f;
      a.sc[i] = i < 2 ? it.sc[e][i] : 1.0f;
    }
  }
  if (e + 1 < it.E) {
    a.vn = 2;
    a.kn = 6;
    for (int i = 0; i < kMaxBasisVars; ++i) {
      a.shn[i] = i < 2 ? it.sh[e + 1][i] : 0.0f;
      a.scn[i] = i < 2 ? it.sc[e + 1][i] : 1.0f;
    }
  }
  return a;
}

// Entries of a product's round-r partials: the 7 x 7 triangle, or the price.
__device__ __forceinline__ int item_round_entries(const LsItem& it, int r) {
  return it.E - 1 - r >= 0 ? tri_count(7) : 1;
}

// Block blk of product it's round r < E (or r == E), given the previous
// round's partials summed into tri (r > 0): the later date's coefficients
// (block 0 stores them), then the path work into the round's slot.  Agent as
// for ls_date_work.
template <bool Agent>
__device__ __forceinline__ void ls_item_round(const LsItem& it, int r, int blk, const double* tri,
                                              double* gram, double* scratch, double* beta,
                                              const PathIn& first, double* lds_d) {
  const DateArgs a = item_round_args(it, r);
  float unused = 0.0f;
  if (r > 0) {
    solve_tri(tri, 6, gram, scratch, beta);
    // the caller's re-arming stores performed before this round's partials
    if (Agent && r >= 2) __builtin_amdgcn_s_waitcnt(0x0F70);
    if (blk == 0 && threadIdx.x < 6)
      gmem(it.coeffs)[static_cast<size_t>(a.e + 1) * 6 + threadIdx.x] = beta[threadIdx.x];
  }
  double* out = gmem(it.xbuf) + ls_slot(r, it.ls_blocks);
  if (a.e >= 0)
    ls_date_work<2, Agent, 8>(a, gmem(it.bv), nullptr, gmem(it.ev), gmem(it.num), beta,
                              gmem(it.Y), false, unused, out, blk, it.ls_blocks, first, lds_d);
  else
    ls_final_work<Agent, 8>(a, gmem(it.bv), nullptr, gmem(it.ev), gmem(it.num), beta, gmem(it.Y),
                            false, 0.0f, out, blk, it.ls_blocks, first);
}

__device__ __forceinline__ PathIn item_first_path(const LsItem& it, int r, int blk) {
  return load_path(item_round_args(it, r), gmem(it.bv), nullptr, gmem(it.ev), gmem(it.num),
                   gmem(it.Y), min(blk * static_cast<int>(blockDim.x) + static_cast<int>(threadIdx.x),
                                   it.P - 1));
}

// Round r of every product (one launch per round; blocks numbered product
// after product).
__global__ void __launch_bounds__(kGramThreads)
    k_ls_round(const LsItem* __restrict__ items, int n, int r) {
  __shared__ double tri[kTriMax];
  __shared__ double lds_d[(kGramThreads / 64) * 256];
  __shared__ double gram[kMaxGram * kMaxGram];
  __shared__ double scratch[kSolveScratch];
  __shared__ double beta[kMaxGram];
  const LsItem& it = items[find_block_owner(items, n, static_cast<int>(blockIdx.x), &LsItem::ls_block0)];
  if (it.E - 1 - r < -1) return;  // this product finished in an earlier round
  const int blk = blockIdx.x - it.ls_block0;
  const PathIn first = item_first_path(it, r, blk);
  if (r > 0)
    sum_tri<false>(gmem(it.xbuf) + ls_slot(r - 1, it.ls_blocks), it.ls_blocks,
                   item_round_entries(it, r - 1), tri, nullptr);
  ls_item_round<false>(it, r, blk, tri, gram, scratch, beta, first, lds_d);
}

// First-pass price of each product (one block per product) from its last
// round's partials.
__global__ void __launch_bounds__(256) k_ls_price_batch(const LsItem* __restrict__ items) {
  const LsItem& it = items[blockIdx.x];
  finalize_col(gmem(it.xbuf) + ls_slot(it.E, it.ls_blocks), it.ls_blocks, 1, 0,
               1.0 / static_cast<double>(it.P), gmem(it.price_out));
}

// Arms every product's three partial slots with kUnsetBits (one block per
// product) before the persistent launch.
__global__ void __launch_bounds__(256) k_ls_portfolio_arm(const LsItem* __restrict__ items) {
  const LsItem& it = items[blockIdx.x];
  unsigned long long* x = reinterpret_cast<unsigned long long*>(gmem(it.xbuf));
  const size_t words = 3 * static_cast<size_t>(kTriMax) * it.ls_blocks;
  for (size_t i = threadIdx.x; i < words; i += blockDim.x) x[i] = kUnsetBits;
}

// Every LS first pass of the portfolio in one persistent launch.  The
// resident blocks share out the products' LS blocks (block b takes LS blocks
// b, b + gridDim.x, ...) and walk the rounds; round E_q + 1 is product q's
// price, summed by its LS block 0.  In round r an LS block of product q
// needs every column of q's round r - 1 (armed words, as in k_ls_first_pass);
// a block tries its LS blocks of the round in turn (one read of the partials
// each), works on each as soon as its product is ready, and sleeps only when
// none was, so products wait only for themselves.  Every LS block of round r
// - 1 belongs to a block that is in round r - 1 or later, and a block in the
// lowest round waits only on finished work, so every wait ends.  Re-arming
// as in k_ls_first_pass.  Same per-block work and solves as the per-round
// launches.  If a wait gives up: NaN prices and sync[0] = 1.
__global__ void __launch_bounds__(kGramThreads)
    k_ls_portfolio(const LsItem* __restrict__ items, int n, int total, int max_ex,
                   unsigned* __restrict__ sync) {
  __shared__ int flag;
  __shared__ double tri[kTriMax];
  __shared__ double lds_d[(kGramThreads / 64) * 256];
  __shared__ double gram[kMaxGram * kMaxGram];
  __shared__ double scratch[kSolveScratch];
  __shared__ double beta[kMaxGram];
  const int g = gridDim.x;
  const int mine = (total - static_cast<int>(blockIdx.x) + g - 1) / g;
  for (int r = 0; r <= max_ex + 1; ++r) {
    unsigned polls = 0;
    // the block's LS blocks of this round, tracked 32 at a time
    for (int i0 = 0; i0 < mine; i0 += 32) {
      const int cnt = mine - i0 < 32 ? mine - i0 : 32;
      unsigned todo = cnt == 32 ? 0xFFFFFFFFu : ((1u << cnt) - 1u);
      while (todo) {
        bool progressed = false;
        for (int i = 0; i < cnt; ++i) {
          if (!((todo >> i) & 1u)) continue;
          const int b = blockIdx.x + (i0 + i) * g;
          const LsItem& it = items[find_block_owner(items, n, b, &LsItem::ls_block0)];
          const int blk = b - it.ls_block0;
          const int nb = it.ls_blocks;
          g_double* xb = it.xbuf;
          bool skip = r > it.E + (blk == 0 ? 1 : 0);  // finished
          if (!skip && r == it.E + 1) {
            // q's price: its LS block 0, once every price partial is in
            double* out = gmem(xb) + ls_slot(it.E, nb);
            if (!sum_tri<true>(out, nb, 1, tri, &flag, 1)) continue;
            finalize_col<true>(out, nb, 1, 0, 1.0 / static_cast<double>(it.P),
                               gmem(it.price_out));
          } else if (!skip) {
            if (r > 0 && !sum_tri<true>(gmem(xb) + ls_slot(r - 1, nb), nb,
                                        item_round_entries(it, r - 1), tri, &flag, 1))
              continue;
            const PathIn first = item_first_path(it, r, blk);
            if (r >= 2) rearm_tri(gmem(xb) + ls_slot(r - 2, nb), nb, item_round_entries(it, r - 2), blk);
            ls_item_round<true>(it, r, blk, tri, gram, scratch, beta, first, lds_d);
            __syncthreads();  // LDS reused by the next LS block
          }
          todo &= ~(1u << i);
          progressed = true;
        }
        if (!progressed) {
          __builtin_amdgcn_s_sleep(4);
          if (++polls > kBarrierPolls) goto fail;
        }
      }
    }
  }
  return;
fail:
  if (threadIdx.x == 0) {
    sync[0] = 1u;
    for (int q = 0; q < n; ++q) *gmem(items[q].price_out) = __builtin_nan("");
  }
}

}  // namespace
}  // namespace koo

namespace koo {
size_t ls_workspace_xbuf_offset(int paths) { return ls_layout(paths).xbuf; }
int ls_workspace_xbuf_columns(int paths) { return gram_blocks(paths); }
}  // namespace koo

using namespace koo;

// ---------------------------------------------------------------------------
// Fused Bermudan path quantities.
// ---------------------------------------------------------------------------
extern "C" int bermudan_paths_gpu(int paths, int rates, int factors, const float* variates,
                                  const float* initial_rates, const float* taus,
                                  const float* displacements, const float* pseudo_roots,
                                  int n_ex, const int* exercise_steps, float strike,
                                  float* basis_variables, float* exercise_values,
                                  float* numeraires, void* stream) {
  if (paths < 1 || rates < 1 || rates > kMaxRates || factors < 1 || factors > kMaxFactors ||
      n_ex < 1 || n_ex > rates)
    return set_error(KOO_ERR_INVALID,
                     "bermudan_paths_gpu: need paths>=1, 1<=rates<=%d, 1<=factors<=%d, "
                     "1<=exercise dates<=rates (got %d, %d, %d, %d)",
                     kMaxRates, kMaxFactors, paths, rates, factors, n_ex);
  if (!variates || !initial_rates || !taus || !displacements || !pseudo_roots || !exercise_steps ||
      !basis_variables || !exercise_values || !numeraires)
    return set_error(KOO_ERR_INVALID, "bermudan_paths_gpu: null pointer");
  ExMap ex{};
  for (int s = 0; s < kMaxRates; ++s) ex.e_of_step[s] = -1;
  for (int e = 0; e < n_ex; ++e) {
    const int s = exercise_steps[e];
    if (s < 0 || s >= rates || (e > 0 && s <= exercise_steps[e - 1]))
      return set_error(KOO_ERR_INVALID,
                       "bermudan_paths_gpu: exercise steps must increase within [0, %d)", rates);
    ex.e_of_step[s] = static_cast<signed char>(e);
  }
  ex.last = exercise_steps[n_ex - 1];
  const int threads = berm_threads(rates);
  const int blocks = (paths + threads - 1) / threads;
  const size_t lds = sizeof(float) * (static_cast<size_t>(rates) * threads + step_table_floats(rates));
  hipStream_t st = as_stream(stream);
  switch (factors) {
#define KOO_CASE(F)                                                                             \
  case F:                                                                                       \
    hipLaunchKernelGGL(k_berm_paths<F>, dim3(blocks), dim3(threads), lds, st, variates,         \
                       initial_rates, taus, displacements, pseudo_roots, strike, paths, rates, 1, \
                       ex, basis_variables, exercise_values, numeraires);                       \
    break;
    KOO_CASE(1) KOO_CASE(2) KOO_CASE(3) KOO_CASE(4) KOO_CASE(5) KOO_CASE(6) KOO_CASE(7)
    KOO_CASE(8)
#undef KOO_CASE
  }
  return check_launch("bermudan_paths_gpu");
}

// ---------------------------------------------------------------------------
// Fused LS first pass, frozen-strategy pass, normal equations, Cholesky.
// ---------------------------------------------------------------------------
extern "C" int ls_workspace(int paths, size_t* bytes) {
  if (paths < 1) return set_error(KOO_ERR_INVALID, "ls_workspace: paths must be >= 1");
  if (bytes) *bytes = ls_layout(paths).total;
  return KOO_OK;
}

extern "C" int ls_first_pass_gpu(int n_ex, const float* basis_variables, const int* vars_each_step,
                                 int max_vars, const float* cash_flows, const float* exercise_values,
                                 const float* numeraires, int paths, int use_cross, int itm_only,
                                 const float* var_shift, const float* var_scale, void* workspace,
                                 double* coefficients, double* price_out, void* stream) {
  if (int e = check_ls("ls_first_pass_gpu", n_ex, vars_each_step, max_vars, paths, use_cross))
    return e;
  if (!basis_variables || !exercise_values || !numeraires || !workspace || !coefficients ||
      !price_out)
    return set_error(KOO_ERR_INVALID, "ls_first_pass_gpu: null pointer");
  const LsLayout lay = ls_layout(paths);
  char* ws = static_cast<char*>(workspace);
  float* Y = reinterpret_cast<float*>(ws + lay.y);
  double* xbuf = reinterpret_cast<double*>(ws + lay.xbuf);
  hipStream_t st = as_stream(stream);
  const bool cross = use_cross != 0;
  AllDates d{};
  d.E = n_ex;
  d.P = paths;
  d.max_vars = max_vars;
  d.maxd = basis_count(max_vars, cross);
  d.cross = cross;
  d.itm_only = itm_only != 0;
  bool small = true;  // the persistent kernel takes bases of at most 7 functions
  for (int e = 0; e < n_ex; ++e) {
    d.v[e] = static_cast<signed char>(vars_each_step[e]);
    d.k[e] = static_cast<signed char>(basis_count(vars_each_step[e], cross));
    fill_norm(var_shift, var_scale, e, vars_each_step[e], max_vars, d.sh[e], d.sc[e]);
    small = small && d.k[e] + 1 <= 8;
  }
  bool resident = false;
  const int nb = ls_pass_blocks(paths, &resident);
  // One launch per round by default: on MI355X at 2^18 paths, 10 dates, the
  // launch boundary exchanges the partials faster than the persistent
  // kernel's polling exchange (141 against 171 us per pass).  One persistent
  // launch over all rounds (after arming its partial words) with
  // KOO_LS_LAUNCH_PER_DATE=0, when every basis has at most 7 functions and
  // the blocks can all be resident.
  const char* per_date = getenv("KOO_LS_LAUNCH_PER_DATE");
  if (small && resident && per_date && per_date[0] == '0') {
    unsigned* bar = reinterpret_cast<unsigned*>(ws + lay.bar);
    const size_t words = 3 * static_cast<size_t>(kTriMax) * nb;
    hipLaunchKernelGGL(k_fill_u64, dim3((words + 255) / 256), dim3(256), 0, st,
                       reinterpret_cast<unsigned long long*>(xbuf), words, kUnsetBits);
    hipLaunchKernelGGL(k_zero_words, dim3(1), dim3(64), 0, st, bar, 1);
    hipLaunchKernelGGL(k_ls_first_pass, dim3(nb), dim3(kGramThreads), 0, st, d, basis_variables,
                       cash_flows, exercise_values, numeraires, coefficients, Y, xbuf, bar,
                       price_out);
    return check_launch("ls_first_pass_gpu");
  }
  for (int r = 0; r <= n_ex; ++r) {
    const int e = n_ex - 1 - r;
    if (e < 0 || d.k[e] + 1 <= 8)
      hipLaunchKernelGGL(k_ls_round_single<2>, dim3(nb), dim3(kGramThreads), 0, st, d, r,
                         basis_variables, cash_flows, exercise_values, numeraires, coefficients,
                         Y, xbuf);
    else
      hipLaunchKernelGGL(k_ls_round_single<1>, dim3(nb), dim3(kGramThreads), 0, st, d, r,
                         basis_variables, cash_flows, exercise_values, numeraires, coefficients,
                         Y, xbuf);
  }
  hipLaunchKernelGGL(k_finalize_cols, dim3(1), dim3(256), 0, st, xbuf + ls_slot(n_ex, nb), nb, 1,
                     1.0 / static_cast<double>(paths), price_out);
  return check_launch("ls_first_pass_gpu");
}

extern "C" int ls_second_pass_gpu(int n_ex, const float* basis_variables, const int* vars_each_step,
                                  int max_vars, const float* cash_flows,
                                  const float* exercise_values, const float* numeraires, int paths,
                                  int use_cross, int itm_only, const float* var_shift,
                                  const float* var_scale, const double* coefficients,
                                  void* workspace, double* price_out, void* stream) {
  if (int e = check_ls("ls_second_pass_gpu", n_ex, vars_each_step, max_vars, paths, use_cross))
    return e;
  if (!basis_variables || !exercise_values || !numeraires || !workspace || !coefficients ||
      !price_out)
    return set_error(KOO_ERR_INVALID, "ls_second_pass_gpu: null pointer");
  const LsLayout lay = ls_layout(paths);
  double* part = reinterpret_cast<double*>(static_cast<char*>(workspace) + lay.part);
  AllDates d{};
  d.E = n_ex;
  d.P = paths;
  d.max_vars = max_vars;
  d.maxd = basis_count(max_vars, use_cross != 0);
  d.cross = use_cross != 0;
  d.itm_only = itm_only != 0;
  for (int e = 0; e < n_ex; ++e) {
    d.v[e] = static_cast<signed char>(vars_each_step[e]);
    d.k[e] = static_cast<signed char>(basis_count(vars_each_step[e], use_cross != 0));
    fill_norm(var_shift, var_scale, e, vars_each_step[e], max_vars, d.sh[e], d.sc[e]);
  }
  hipStream_t st = as_stream(stream);
  // one lane per path up to 8 blocks per CU (the workspace's partial region
  // holds kMaxGram^2 doubles per ls_blocks block, more than these blocks)
  const int nb = static_cast<int>(grid_for(paths, kLsThreads, 8 * device_cu_count()));
  hipLaunchKernelGGL(k_ls_second, dim3(nb), dim3(kLsThreads), 0, st, d, basis_variables,
                     cash_flows, exercise_values, numeraires, coefficients, part);
  hipLaunchKernelGGL(k_finalize_cols, dim3(1), dim3(256), 0, st, part, nb, 1,
                     1.0 / static_cast<double>(paths), price_out);
  return check_launch("ls_second_pass_gpu");
}

extern "C" int ls_normal_eq_gpu(int paths, int n_basis, const float* basis_functions,
                                const float* target, const float* mask, void* workspace,
                                double* gram_out, void* stream) {
  const int n = n_basis + (target ? 1 : 0);
  if (paths < 1 || n_basis < 1 || n > kMaxGram)
    return set_error(KOO_ERR_INVALID,
                     "ls_normal_eq_gpu: need paths>=1 and 1<=rows<=%d (got %d, %d)", kMaxGram,
                     paths, n);
  if (!basis_functions || !workspace || !gram_out)
    return set_error(KOO_ERR_INVALID, "ls_normal_eq_gpu: null pointer");
  const float* rows[kMaxGram];
  for (int i = 0; i < n_basis; ++i) rows[i] = basis_functions + static_cast<size_t>(i) * paths;
  if (target) rows[n_basis] = target;
  const LsLayout lay = ls_layout(paths);
  double* part = reinterpret_cast<double*>(static_cast<char*>(workspace) + lay.part);
  return gram_launch("ls_normal_eq_gpu", rows, n, mask, paths, part, gram_blocks(paths), 0, 0, n, n,
                     n, gram_out, nullptr, 1.0f, 0.0f, 0, as_stream(stream));
}

namespace {
__global__ void k_cholesky(int n, const double* __restrict__ A, const double* __restrict__ b,
                           double* __restrict__ x, int* __restrict__ info) {
  __shared__ double G[kMaxGram * kMaxGram];
  __shared__ double S[kMaxGram * kMaxGram + 2 * kMaxGram];
  const int m = n + 1;
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j < n; ++j) G[i * m + j] = A[i * n + j];
    G[i * m + n] = b[i];
  }
  double* beta = S + kMaxGram * kMaxGram + kMaxGram;
  const bool ok = cholesky_rhs(G, m, n, 1.0, beta, S, S + kMaxGram * kMaxGram);
  for (int i = 0; i < n; ++i) x[i] = beta[i];
  if (info) *info = ok ? 0 : 1;
}
}  // namespace

extern "C" int cholesky_solve_f64_gpu(int n, const double* A, const double* b, double* x,
                                      int* info, void* stream) {
  if (n < 1 || n >= kMaxGram)
    return set_error(KOO_ERR_INVALID, "cholesky_solve_f64_gpu: need 1<=n<%d (got %d)", kMaxGram,
                     n);
  if (!A || !b || !x) return set_error(KOO_ERR_INVALID, "cholesky_solve_f64_gpu: null pointer");
  hipLaunchKernelGGL(k_cholesky, dim3(1), dim3(1), 0, as_stream(stream), n, A, b, x, info);
  return check_launch("cholesky_solve_f64_gpu");
}

// ---------------------------------------------------------------------------
// The reference's LS launchers.
// ---------------------------------------------------------------------------
extern "C" int adjoinBasisVariablesCaller_gpu(int use_logs, const int* integer_data,
                                              const float* float_data, const float* forwards,
                                              const float* discount_ratios, const float* rate1,
                                              const float* rate2, const float* rate3,
                                              int paths_per_batch, int total_paths,
                                              int paths_previously_done, int steps_and_rates,
                                              int number_exercise_times,
                                              const int* exercise_indices,
                                              const int* exercise_indicators,
                                              float* basis_variables, void* stream) {
  (void)integer_data;
  (void)forwards;
  (void)discount_ratios;
  (void)steps_and_rates;
  (void)exercise_indices;
  (void)exercise_indicators;
  if (paths_per_batch < 1 || total_paths < paths_per_batch + paths_previously_done ||
      paths_previously_done < 0 || number_exercise_times < 1)
    return set_error(KOO_ERR_INVALID,
                     "adjoinBasisVariablesCaller_gpu: bad sizes (batch %d, total %d, done %d, "
                     "exercise times %d)",
                     paths_per_batch, total_paths, paths_previously_done, number_exercise_times);
  if (!rate1 || !rate2 || !rate3 || !basis_variables || (use_logs && !float_data))
    return set_error(KOO_ERR_INVALID, "adjoinBasisVariablesCaller_gpu: null pointer");
  hipLaunchKernelGGL(k_basis_variables, dim3(grid_for(paths_per_batch, 256)), dim3(256), 0,
                     as_stream(stream), use_logs, float_data, rate1, rate2, rate3, paths_per_batch,
                     total_paths, paths_previously_done, number_exercise_times, basis_variables);
  return check_launch("adjoinBasisVariablesCaller_gpu");
}

extern "C" int adjoinExerciseValues_Bermudan_swaption_gpu(
    float* exercise_values, const float* aux_float_data, const int* aux_int_data,
    const int* exercise_indices, const int* exercise_indicators, int number_exercise_times,
    int paths_for_output, int output_offset, int batch_paths, int number_steps_and_rates,
    const float* rates1, const float* rates2, const float* rates3, const float* forwards,
    const float* disc_ratios, void* stream) {
  (void)aux_int_data;
  (void)exercise_indices;
  (void)exercise_indicators;
  (void)number_steps_and_rates;
  (void)rates3;
  (void)forwards;
  (void)disc_ratios;
  if (batch_paths < 1 || output_offset < 0 || paths_for_output < batch_paths + output_offset ||
      number_exercise_times < 1)
    return set_error(KOO_ERR_INVALID,
                     "adjoinExerciseValues_Bermudan_swaption_gpu: bad sizes (batch %d, output "
                     "%d, offset %d, exercise times %d)",
                     batch_paths, paths_for_output, output_offset, number_exercise_times);
  if (!exercise_values || !aux_float_data || !rates1 || !rates2)
    return set_error(KOO_ERR_INVALID, "adjoinExerciseValues_Bermudan_swaption_gpu: null pointer");
  hipLaunchKernelGGL(k_exercise_values_bermudan, dim3(grid_for(batch_paths, 256)), dim3(256), 0,
                     as_stream(stream), aux_float_data, rates1, rates2, batch_paths,
                     paths_for_output, output_offset, number_exercise_times, exercise_values);
  return check_launch("adjoinExerciseValues_Bermudan_swaption_gpu");
}

extern "C" int expandBasisFunctions_quadratic_gpu(int total_paths, const float* basis_variables,
                                                  int exercise_number, int max_basis_variable_size,
                                                  int variable_size_this_step, int use_cross_terms,
                                                  float* basis_functions, void* stream) {
  if (total_paths < 1 || exercise_number < 0 || variable_size_this_step < 0 ||
      variable_size_this_step > max_basis_variable_size ||
      variable_size_this_step > kMaxBasisVars)
    return set_error(KOO_ERR_INVALID,
                     "expandBasisFunctions_quadratic_gpu: need paths>=1, 0<=variables<=min(max "
                     "size, %d) (got %d, %d, %d)",
                     kMaxBasisVars, total_paths, variable_size_this_step, max_basis_variable_size);
  if (!basis_variables || !basis_functions)
    return set_error(KOO_ERR_INVALID, "expandBasisFunctions_quadratic_gpu: null pointer");
  const float* in = basis_variables + static_cast<size_t>(total_paths) * max_basis_variable_size *
                                          exercise_number;
  hipLaunchKernelGGL(k_expand_basis, dim3(grid_for(total_paths, 256)), dim3(256), 0,
                     as_stream(stream), in, total_paths, variable_size_this_step, use_cross_terms,
                     basis_functions);
  return check_launch("expandBasisFunctions_quadratic_gpu");
}

extern "C" int continuation_net_value_evaluation_gpu(int paths, int exercise_number,
                                                     const float* basis_functions,
                                                     int number_basis_functions,
                                                     const float* basis_weights,
                                                     const float* exercise_values,
                                                     float andersen_shift, float* output,
                                                     void* stream) {
  (void)exercise_number;
  if (paths < 1 || number_basis_functions < 1)
    return set_error(KOO_ERR_INVALID, "continuation_net_value_evaluation_gpu: bad sizes");
  if (!basis_functions || !basis_weights || !exercise_values || !output)
    return set_error(KOO_ERR_INVALID, "continuation_net_value_evaluation_gpu: null pointer");
  hipLaunchKernelGGL(k_net_continuation, dim3(grid_for(paths, 256)), dim3(256), 0,
                     as_stream(stream), paths, number_basis_functions, basis_weights,
                     basis_functions, exercise_values, andersen_shift, output);
  return check_launch("continuation_net_value_evaluation_gpu");
}

extern "C" int oneStepUpdateLS_multi_gpu(int paths, int exercise_number,
                                         const float* basis_functions, int number_basis_functions,
                                         int basis_functions_row_size, int max_regression_depth,
                                         const float* basis_weights, const float* lower_cuts,
                                         const float* upper_cuts, float* continuation_values,
                                         const float* deflated_cash_flows,
                                         const float* exercise_values, float andersen_shift,
                                         float* est_continuation_values, int* exercises,
                                         void* stream) {
  (void)deflated_cash_flows;
  if (paths < 1 || exercise_number < 0 || number_basis_functions < 1 ||
      basis_functions_row_size < number_basis_functions || max_regression_depth < 1)
    return set_error(KOO_ERR_INVALID, "oneStepUpdateLS_multi_gpu: bad sizes");
  if (!basis_functions || !basis_weights || !lower_cuts || !upper_cuts || !continuation_values ||
      !exercise_values || !est_continuation_values || !exercises)
    return set_error(KOO_ERR_INVALID, "oneStepUpdateLS_multi_gpu: null pointer");
  const float* w = basis_weights + static_cast<size_t>(basis_functions_row_size) *
                                       exercise_number * max_regression_depth;
  const float* ev = exercise_values + static_cast<size_t>(paths) * exercise_number;
  hipLaunchKernelGGL(k_update_multi, dim3(grid_for(paths, 256)), dim3(256), 0, as_stream(stream),
                     paths, number_basis_functions, basis_functions_row_size,
                     max_regression_depth, w, lower_cuts, upper_cuts, basis_functions,
                     continuation_values, ev, andersen_shift, est_continuation_values, exercises);
  return check_launch("oneStepUpdateLS_multi_gpu");
}

extern "C" int updateDeflationOfContinuationValues_gpu(float* deflated_next_step_values,
                                                       const float* numeraire_values, int paths,
                                                       int new_step_number, int old_step_number,
                                                       void* stream) {
  if (paths < 1 || new_step_number < 0 || old_step_number < 0)
    return set_error(KOO_ERR_INVALID, "updateDeflationOfContinuationValues_gpu: bad sizes");
  if (!deflated_next_step_values || !numeraire_values)
    return set_error(KOO_ERR_INVALID, "updateDeflationOfContinuationValues_gpu: null pointer");
  const size_t P = static_cast<size_t>(paths);
  hipLaunchKernelGGL(k_deflate, dim3(grid_for(paths, 256)), dim3(256), 0, as_stream(stream),
                     deflated_next_step_values, numeraire_values + new_step_number * P,
                     numeraire_values + old_step_number * P, paths);
  return check_launch("updateDeflationOfContinuationValues_gpu");
}

// ---------------------------------------------------------------------------
// Outer products (outerProduct_gpu.h:12-92) on the MFMA Gram kernel: blocks of
// at most 8 rows of each operand are stacked into one Gram matrix whose
// off-diagonal block is the product.  The FP64 block partials live in the
// caller's float workspace: the reference launchers keep the reference's
// contract (at least blocks * row_size * col_size floats) and the number of
// partial blocks is sized to fit it; the cuBLAS-named variants take a
// workspace of outer_product_workspace_floats(paths) floats.
// ---------------------------------------------------------------------------
namespace {

int outer_blocks(int paths) { return gram_blocks(paths); }

int outer_product(const char* who, int paths, int R, int C, bool symmetric, const float* in1,
                  const float* in2, float* answer, float* workspace, size_t workspace_floats,
                  float alpha, float beta, bool column_major, hipStream_t st) {
  if (paths < 1 || R < 1 || C < 1)
    return set_error(KOO_ERR_INVALID, "%s: need paths, rows, columns >= 1 (got %d, %d, %d)", who,
                     paths, R, C);
  if (!in1 || (!symmetric && !in2) || !answer || !workspace)
    return set_error(KOO_ERR_INVALID, "%s: null pointer", who);
  if (reinterpret_cast<uintptr_t>(workspace) % alignof(double) != 0)
    return set_error(KOO_ERR_INVALID, "%s: workspace must be 8-byte aligned", who);
  double* part = reinterpret_cast<double*>(workspace);
  const size_t P = static_cast<size_t>(paths);
  const float* rows[kMaxGram];
  // answer(i, j) = sum_p in1[i][p] in2[j][p]; row-major [R][C] (reference
  // kernels) or column-major with leading dimension R (the cuBLAS variants).
  const int ld = column_major ? R : C;
  for (int r0 = 0; r0 < R; r0 += 8) {
    const int nr = R - r0 < 8 ? R - r0 : 8;
    for (int c0 = 0; c0 < C; c0 += 8) {
      const int nc = C - c0 < 8 ? C - c0 : 8;
      for (int i = 0; i < nr; ++i) rows[i] = in1 + (r0 + i) * P;
      const float* second = symmetric ? in1 : in2;
      for (int j = 0; j < nc; ++j) rows[nr + j] = second + (c0 + j) * P;
      float* out = answer + (column_major ? static_cast<size_t>(c0) * ld + r0
                                          : static_cast<size_t>(r0) * ld + c0);
      const int n = nr + nc;
      const size_t fit = workspace_floats * sizeof(float) / (sizeof(double) * n * n);
      const int nb = static_cast<int>(fit < static_cast<size_t>(outer_blocks(paths))
                                          ? fit
                                          : static_cast<size_t>(outer_blocks(paths)));
      if (nb < 1)
        return set_error(KOO_ERR_INVALID, "%s: workspace of %zu floats holds no %dx%d partial",
                         who, workspace_floats, n, n);
      if (int e = gram_launch(who, rows, n, nullptr, paths, part, nb, 0, nr, nr, nc, ld,
                              nullptr, out, alpha, beta, column_major ? 1 : 0, st))
        return e;
    }
  }
  return KOO_OK;
}

}  // namespace

extern "C" size_t outer_product_workspace_floats(int paths) {
  return sizeof(double) / sizeof(float) * static_cast<size_t>(kMaxGram) * kMaxGram *
         outer_blocks(paths < 1 ? 1 : paths);
}

namespace {
size_t reference_workspace(int blocks, int R, int C) {
  return static_cast<size_t>(blocks > 0 ? blocks : 1) * (R > 0 ? R : 1) * (C > 0 ? C : 1);
}
}  // namespace

extern "C" int reduceOuterProduct_gpu(int paths, int row_size, int col_size, int symmetric,
                                      int threads, int blocks, const float* input_data_global1,
                                      const float* input_data_global2, float* answer_global,
                                      float* workspace_data_global, void* stream) {
  (void)threads;
  const int C = symmetric ? row_size : col_size;
  return outer_product("reduceOuterProduct_gpu", paths, row_size, C, symmetric != 0,
                       input_data_global1, input_data_global2, answer_global,
                       workspace_data_global, reference_workspace(blocks, row_size, C), 1.0f,
                       0.0f, false, as_stream(stream));
}

extern "C" int reducedOuterProduct_gpu(int paths, int row_size, int col_size, int threads,
                                       int blocks, const float* input_data_global1,
                                       const float* input_data_global2, float* answer_global,
                                       float* workspace_data_global, void* stream) {
  (void)threads;
  return outer_product("reducedOuterProduct_gpu", paths, row_size, col_size, false,
                       input_data_global1, input_data_global2, answer_global,
                       workspace_data_global, reference_workspace(blocks, row_size, col_size),
                       1.0f, 0.0f, false, as_stream(stream));
}

extern "C" int reduceOuterProductSymmetric_gpu(int paths, int row_size, int threads, int blocks,
                                               const float* input_data_global1,
                                               float* answer_global, float* workspace_data_global,
                                               void* stream) {
  (void)threads;
  return outer_product("reduceOuterProductSymmetric_gpu", paths, row_size, row_size, true,
                       input_data_global1, nullptr, answer_global, workspace_data_global,
                       reference_workspace(blocks, row_size, row_size), 1.0f, 0.0f, false,
                       as_stream(stream));
}

extern "C" int reduceOuterProductSymmetric_cublas(int paths, int row_size, float alpha, float beta,
                                                  const float* input_data_global,
                                                  float* answer_global,
                                                  float* workspace_data_global, void* stream) {
  return outer_product("reduceOuterProductSymmetric_cublas", paths, row_size, row_size, true,
                       input_data_global, nullptr, answer_global, workspace_data_global,
                       outer_product_workspace_floats(paths), alpha, beta, true,
                       as_stream(stream));
}

extern "C" int reduceOuterProduct_cublas(int paths, int row_size, int col_size, float alpha,
                                         float beta, const float* input_data_global1,
                                         const float* input_data_global2, float* answer_global,
                                         float* workspace_data_global, void* stream) {
  return outer_product("reduceOuterProduct_cublas", paths, row_size, col_size, false,
                       input_data_global1, input_data_global2, answer_global,
                       workspace_data_global, outer_product_workspace_floats(paths), alpha, beta,
                       true, as_stream(stream));
}

// ---------------------------------------------------------------------------
// Batched portfolio: path quantities and LS first passes.
// ---------------------------------------------------------------------------
// The descriptor table, then (at the next 256-byte boundary) the persistent
// LS kernel's sync words: two per product (LS blocks done, solves published)
// and an error word.
size_t ls_portfolio_bar_offset(int n_products) {
  return align256(sizeof(LsItem) * static_cast<size_t>(n_products));
}
int ls_portfolio_sync_words(int) { return 1; }  // the error word

extern "C" size_t bermudan_ls_portfolio_descriptor_bytes(int n_products) {
  return n_products > 0 ? ls_portfolio_bar_offset(n_products) +
                              align256(sizeof(unsigned) * ls_portfolio_sync_words(n_products))
                        : 0;
}

extern "C" int bermudan_ls_portfolio_prepare(int n_products, const koo_ls_product* products,
                                             void* descriptors_device, koo_portfolio_plan* plan) {
  const char* who = "bermudan_ls_portfolio_prepare";
  if (n_products < 1 || !products || !descriptors_device || !plan)
    return set_error(KOO_ERR_INVALID, "%s: bad arguments", who);
  std::vector<LsItem> items(n_products);
  std::vector<double> cost(n_products);
  int max_n = 0, max_e = 0;
  for (int i = 0; i < n_products; ++i) {
    const koo_ls_product& q = products[i];
    if (q.paths < 1 || q.rates < 1 || q.rates > 50 || q.factors < 1 || q.factors > 5 ||
        q.n_ex < 1 || q.n_ex > kLsBatchMaxDates || q.n_ex > q.rates)
      return set_error(KOO_ERR_INVALID,
                       "%s: product %d: need paths>=1, 1<=rates<=50, 1<=factors<=5, "
                       "1<=dates<=min(rates,%d)",
                       who, i, kLsBatchMaxDates);
    if (!q.variates || !q.initial_rates || !q.taus || !q.displacements || !q.pseudo_roots ||
        !q.exercise_steps || !q.basis_variables || !q.exercise_values || !q.numeraires ||
        !q.workspace || !q.coefficients || !q.price_out)
      return set_error(KOO_ERR_INVALID, "%s: product %d: null pointer", who, i);
    LsItem& it = items[i];
    it = LsItem{};
    for (int s = 0; s < kMaxRates; ++s) it.ex.e_of_step[s] = -1;
    for (int e = 0; e < q.n_ex; ++e) {
      const int s = q.exercise_steps[e];
      if (s < 0 || s >= q.rates || (e > 0 && s <= q.exercise_steps[e - 1]))
        return set_error(KOO_ERR_INVALID, "%s: product %d: exercise steps must increase", who, i);
      it.ex.e_of_step[s] = static_cast<signed char>(e);
      for (int v = 0; v < 2; ++v) {
        it.sh[e][v] = q.var_shift ? q.var_shift[2 * e + v] : 0.0f;
        it.sc[e][v] = q.var_scale ? q.var_scale[2 * e + v] : 1.0f;
      }
    }
    it.ex.last = q.exercise_steps[q.n_ex - 1];
    const LsLayout lay = ls_layout(q.paths);
    char* ws = static_cast<char*>(q.workspace);
    it.z = (g_cfloat*)q.variates;
    it.f0 = (g_cfloat*)q.initial_rates;
    it.taus = (g_cfloat*)q.taus;
    it.disp = (g_cfloat*)q.displacements;
    it.A = (g_cfloat*)q.pseudo_roots;
    it.bv = (g_float*)q.basis_variables;
    it.ev = (g_float*)q.exercise_values;
    it.num = (g_float*)q.numeraires;
    it.Y = (g_float*)(ws + lay.y);
    it.xbuf = (g_double*)(ws + lay.xbuf);
    it.coeffs = (g_double*)q.coefficients;
    it.price_out = (g_double*)q.price_out;
    it.strike = q.strike;
    it.P = q.paths;
    it.N = q.rates;
    it.F = q.factors;
    it.E = q.n_ex;
    it.itm_only = q.itm_only != 0;
    it.paths_blocks = (q.paths + 255) / 256;
    // LS blocks per product: about two 1024-lane blocks per CU over the whole
    // portfolio (each block also sums the later date's partials and solves
    // its normal equations, so fewer, fuller blocks amortise that)
    {
      const int share = q.ls_blocks > 0 ? q.ls_blocks
                                        : (2 * device_cu_count() + n_products - 1) / n_products;
      const int g = gram_blocks(q.paths);
      it.ls_blocks = share < 1 ? 1 : (share < g ? share : g);
    }
    double u = 0.0;
    for (int s = 0; s <= it.ex.last; ++s) u += q.rates - s;
    cost[i] = u * q.paths * (q.factors + 4);
    if (q.rates > max_n) max_n = q.rates;
    if (q.n_ex > max_e) max_e = q.n_ex;
  }
  // path kernel: rate class first (one launch per class), then most
  // expensive first; LS blocks in the same order
  std::vector<int> order(n_products);
  for (int i = 0; i < n_products; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    const int ca = rate_class(items[a].N), cb = rate_class(items[b].N);
    return ca != cb ? ca < cb : cost[a] > cost[b];
  });
  std::vector<LsItem> sorted(n_products);
  int64_t pb = 0, lb = 0;
  *plan = koo_portfolio_plan{};
  for (int i = 0; i < n_products; ++i) {
    sorted[i] = items[order[i]];
    sorted[i].paths_block0 = static_cast<int>(pb);
    sorted[i].ls_block0 = static_cast<int>(lb);
    const int c = rate_class(sorted[i].N);
    if (plan->classes == 0 || rate_class(sorted[i - 1].N) != c) {
      plan->class_first_block[plan->classes] = static_cast<int>(pb);
      ++plan->classes;
    }
    plan->class_blocks[plan->classes - 1] += sorted[i].paths_blocks;
    if (sorted[i].N > plan->class_max_rates[plan->classes - 1])
      plan->class_max_rates[plan->classes - 1] = sorted[i].N;
    pb += sorted[i].paths_blocks;
    lb += sorted[i].ls_blocks;
  }
  if (pb > 0x7fffffff || lb > 0x7fffffff)
    return set_error(KOO_ERR_INVALID, "%s: too many blocks", who);
  if (hipMemcpy(descriptors_device, sorted.data(), sizeof(LsItem) * n_products,
                hipMemcpyHostToDevice) != hipSuccess)
    return set_error(KOO_ERR_HIP, "%s: descriptor copy failed", who);
  plan->products = n_products;
  plan->blocks = static_cast<int>(pb);
  plan->max_rates = max_n;
  plan->ls_blocks = static_cast<int>(lb);
  plan->max_ex = max_e;
  return KOO_OK;
}

extern "C" int bermudan_ls_portfolio_run(const koo_portfolio_plan* plan,
                                         const void* descriptors_device, void* stream) {
  const char* who = "bermudan_ls_portfolio_run";
  if (!plan || !descriptors_device || plan->products < 1 || plan->blocks < 1 ||
      plan->ls_blocks < 1 || plan->max_rates < 1 || plan->max_rates > 50 || plan->max_ex < 1 ||
      plan->max_ex > kLsBatchMaxDates || plan->classes < 1 || plan->classes > 4)
    return set_error(KOO_ERR_INVALID, "%s: bad plan", who);
  const LsItem* items = static_cast<const LsItem*>(descriptors_device);
  hipStream_t st = as_stream(stream);
  for (int c = 0; c < plan->classes; ++c) {
    const int nc = plan->class_max_rates[c], b0 = plan->class_first_block[c];
    const int nb = plan->class_blocks[c];
    if (nc < 1 || nc > plan->max_rates || nb < 1 || b0 < 0 || b0 + nb > plan->blocks)
      return set_error(KOO_ERR_INVALID, "%s: bad plan class %d", who, c);
    const size_t lds = sizeof(float) * (static_cast<size_t>(nc) * 256 + step_table_floats(nc));
    hipLaunchKernelGGL(k_berm_paths_batch, dim3(nb), dim3(256), lds, st, items, plan->products,
                       b0);
  }
  // The LS first passes: one launch per round and one for the prices (the
  // default: on MI355X the 64-product portfolio's passes take 7.0 ms so
  // against 8.0 ms in one launch), or with KOO_LS_LAUNCH_PER_DATE=0 one
  // persistent launch (after arming every product's partial words and
  // zeroing the error word) with as many blocks as can be resident.
  int per_cu = 0;
  const char* per_date = getenv("KOO_LS_LAUNCH_PER_DATE");
  if (per_date && per_date[0] == '0' &&
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_ls_portfolio, kGramThreads, 0) ==
          hipSuccess &&
      per_cu > 0) {
    const int cap = per_cu * device_cu_count();
    const int grid = plan->ls_blocks < cap ? plan->ls_blocks : cap;
    unsigned* bar = reinterpret_cast<unsigned*>(static_cast<char*>(const_cast<void*>(
                                                    descriptors_device)) +
                                                ls_portfolio_bar_offset(plan->products));
    hipLaunchKernelGGL(k_ls_portfolio_arm, dim3(plan->products), dim3(256), 0, st, items);
    hipLaunchKernelGGL(k_zero_words, dim3(1), dim3(64), 0, st, bar, 1);
    hipLaunchKernelGGL(k_ls_portfolio, dim3(grid), dim3(kGramThreads), 0, st, items,
                       plan->products, plan->ls_blocks, plan->max_ex, bar);
    return check_launch(who);
  }
  for (int r = 0; r <= plan->max_ex; ++r)
    hipLaunchKernelGGL(k_ls_round, dim3(plan->ls_blocks), dim3(kGramThreads), 0, st, items,
                       plan->products, r);
  hipLaunchKernelGGL(k_ls_price_batch, dim3(plan->products), dim3(256), 0, st, items);
  return check_launch(who);
}
