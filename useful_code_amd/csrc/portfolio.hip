// portfolio.hip -- a portfolio of Bermudan payer swaptions priced in one
// persistent launch on gfx950: for every product, the fused Sobol ->
// scramble -> normal -> Brownian bridge generation, the path quantities, the
// Longstaff-Schwartz first pass, the frozen-strategy forward and the adjoint
// backward, then the first-pass price, the frozen price and the 2N Greeks.
//
// The grid is exactly the blocks that can be resident at once; the host
// splits it into one team of blocks per product, sized to the product's
// work.  A team's block t owns the paths of the 256-path chunks t, t + T,
// t + 2T, ... (T the team size) in every phase, and each of its lanes the
// same paths throughout, so every value one phase writes and a later phase
// reads is written and read by the same lane: the phases follow each other
// inside the block with no grid-wide synchronisation.  Only two things
// cross blocks, both within a team: the LS first pass's FP64 partials of each
// round (the sentinel-armed exchange of the single-product persistent pass:
// every block sums all its team's columns and solves the date itself), and
// the final sums (price and Greek partials stored agent-coherent into armed
// words, summed once none is unset).  Every wait is bounded.
//
// The per-block bodies are the single-product kernels' (bridge_device.h,
// ls_device.h, berm_device.h), so given the same variates and coefficients
// the forward and backward are bit for bit those of bermudan_frozen_fwd/bwd.
#include <algorithm>
#include <cstdlib>
#include <vector>

#include "berm_device.h"
#include "bridge_device.h"
#include "ls_device.h"

namespace koo {
namespace {

using namespace lmm;

constexpr int kPfThreads = 256;
constexpr int kPfMaxDates = 40;

struct PfItem {
  // generation (gen == 0: the variates are given)
  int gen, n_scr;
  uint32_t off1;  // Gray index of path 0 (path_offset + 1)
  const uint32_t KOO_GLOBAL* dirs;
  const uint32_t KOO_GLOBAL* scr;
  const int KOO_GLOBAL* alloc;
  const BridgeTables KOO_GLOBAL* tables;
  // model
  g_float* z;
  g_cfloat *f0, *taus, *disp, *A, *vols;
  float strike;
  int P, N, F, k;
  // LS first pass (two variables per date, the cross basis)
  ExMap ex;
  int E, itm_only;
  float sh[kPfMaxDates][2], sc[kPfMaxDates][2];
  g_float *bv, *ev, *num, *Y;
  g_double *xbuf, *coeffs, *price1;
  // frozen strategy forward and adjoint backward
  BermArgs b;
  g_float *ckpt, *lamt, *vout, *scratch;
  g_int* tau;
  g_double *pfwd, *pbwd, *price, *greeks;
  // team
  int team0, team, nvb;
};

__device__ __forceinline__ Model pf_model(const PfItem& it, int k) {
  return Model{gmem(it.z), cmem(it.f0), cmem(it.taus), cmem(it.disp), cmem(it.A), it.strike,
               it.P, it.N, k};
}

// The arguments of round r of the product's first pass (as item_round_args
// of the batched LS: date e = E - 1 - r, e = -1 for round E).
__device__ __forceinline__ DateArgs pf_round_args(const PfItem& it, int r) {
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

__device__ __forceinline__ int pf_round_entries(const PfItem& it, int r) {
  return it.E - 1 - r >= 0 ? tri_count(7) : 1;
}

// Waits until no entry part[b * ncols + col], b < nb, holds kUnsetBits (agent
// loads, bounded); false on every lane if a wait gave up.  Contains barriers.
__device__ __forceinline__ bool wait_column(const double* part, int nb, int ncols, int col,
                                            int* flag_lds) {
  if (threadIdx.x == 0) *flag_lds = 1;
  __syncthreads();
  bool ok = true;
  for (int b0 = 0; ok && b0 < nb; b0 += blockDim.x) {
    const int b = b0 + threadIdx.x;
    unsigned polls = 0;
    for (;;) {
      const bool mine =
          b >= nb || !is_unset(load_agent(part + static_cast<size_t>(b) * ncols + col));
      if (__all(mine)) break;
      if (++polls >= kBarrierPolls) {
        ok = false;
        break;
      }
      __builtin_amdgcn_s_sleep(2);
    }
  }
  if (!ok) *flag_lds = 0;
  __syncthreads();
  ok = *flag_lds != 0;
  __syncthreads();
  return ok;
}

// LDS of the LS phase (doubles, carved from the dynamic buffer): the Gram
// staging (256 per wavefront), the summed triangle, the Gram, the solver's
// scratch and the coefficients.
constexpr int kPfLsDoubles = (kPfThreads / 64) * 256 + kTriMax + kMaxGram * kMaxGram +
                             kSolveScratch + kMaxGram;

// The whole portfolio.  Dynamic LDS: [0, tab_off) floats for the phase in
// progress, then the product's frozen coefficients as FP32 (E x 6), filled
// by the LS rounds and read by the forward.  ticks[block]: the block's run
// time in 100 MHz ticks (what the host rebalances the teams by).  err[0] = 1
// and NaN outputs if a wait gave up.
__global__ void __launch_bounds__(kPfThreads, 3)
    k_portfolio(const PfItem* __restrict__ items, int n, int tab_off, unsigned* __restrict__ err,
                unsigned long long* __restrict__ ticks) {
  extern __shared__ float lds[];
  const unsigned long long t_start = __builtin_amdgcn_s_memrealtime();
  __shared__ int flag;
  const PfItem& it = items[find_block_owner(items, n, static_cast<int>(blockIdx.x), &PfItem::team0)];
  const int t = static_cast<int>(blockIdx.x) - it.team0;
  const int T = it.team;
  const int P = it.P;
  float* tab = lds + tab_off;

  // 1. generation: the bridge variates of the block's paths, factor by factor
  // (the direction numbers of the factor's S dimensions in LDS behind the
  // path storage, as k_sobol_bridge)
  if (it.gen) {
    const int S = it.N;
    uint32_t* v = reinterpret_cast<uint32_t*>(lds + S * kPfThreads);
    for (int f = 0; f < it.F; ++f) {
      const int KOO_GLOBAL* idx = it.alloc + f * S;
      __syncthreads();
      for (int w = threadIdx.x; w < S * 32; w += blockDim.x)
        v[w] = it.dirs[static_cast<size_t>(idx[w >> 5]) * 32 + (w & 31)];
      __syncthreads();
      for (int vb = t; vb < it.nvb; vb += T) {
        const int p = vb * kPfThreads + static_cast<int>(threadIdx.x);
        if (p >= P) continue;  // no barrier inside the bridge
        const uint32_t gi = static_cast<uint32_t>(p) + it.off1;
        bridge_lane(*gmem(it.tables), S, it.F, P, p, f, lds,
                    [&](int i) {
                      const int d = idx[i];
                      const uint32_t s = d < it.n_scr ? it.scr[d] : 0u;
                      return invnorm_u32(sobol_point(v + 32 * i, gi) ^ s);
                    },
                    gmem(it.z));
      }
    }
  }

  // 2. path quantities (basis variables, exercise values, numeraires)
  {
    const Model m = pf_model(it, 1);
    switch (it.F) {
#define KOO_CASE(F)                                                                            \
  case F:                                                                                      \
    for (int vb = t; vb < it.nvb; vb += T) {                                                   \
      __syncthreads();                                                                         \
      berm_paths_block<F>(m, it.ex, gmem(it.bv), gmem(it.ev), gmem(it.num), vb);               \
    }                                                                                          \
    break;
      KOO_CASE(1) KOO_CASE(2) KOO_CASE(3) KOO_CASE(4) KOO_CASE(5)
#undef KOO_CASE
    }
  }
  __syncthreads();

  // 3. LS first pass: rounds r = 0..E over the team's columns
  {
    double* ld = reinterpret_cast<double*>(lds);
    double* lds_d = ld;
    double* tri = lds_d + (kPfThreads / 64) * 256;
    double* gram = tri + kTriMax;
    double* scratch = gram + kMaxGram * kMaxGram;
    double* beta = scratch + kSolveScratch;
    float* Y = gmem(it.Y);
    float yr = 0.0f;
    const bool yin = static_cast<long long>(T) * kPfThreads >= P;
    const float* Yl = yin ? nullptr : Y;
    const int p0 = min(t * kPfThreads + static_cast<int>(threadIdx.x), P - 1);
    double* xbuf = gmem(it.xbuf);
    PathIn first = load_path(pf_round_args(it, 0), gmem(it.bv), nullptr, gmem(it.ev),
                             gmem(it.num), Yl, p0);
    for (int r = 0; r <= it.E; ++r) {
      const DateArgs a = pf_round_args(it, r);
      if (r > 0) {
        if (!sum_tri<true>(xbuf + ls_slot(r - 1, T), T, pf_round_entries(it, r - 1), tri, &flag))
          goto fail;
        if (r >= 2) rearm_tri(xbuf + ls_slot(r - 2, T), T, pf_round_entries(it, r - 2), t);
        solve_tri(tri, 6, gram, scratch, beta);
        if (r >= 2) __builtin_amdgcn_s_waitcnt(0x0F70);  // re-arming performed first
        if (threadIdx.x < 6) {
          tab[(a.e + 1) * 6 + threadIdx.x] = static_cast<float>(beta[threadIdx.x]);
          if (t == 0) gmem(it.coeffs)[(a.e + 1) * 6 + threadIdx.x] = beta[threadIdx.x];
        }
      }
      double* out = xbuf + ls_slot(r, T);
      if (a.e >= 0) {
        ls_date_work<2, true, 8>(a, gmem(it.bv), nullptr, gmem(it.ev), gmem(it.num), beta, Y, yin,
                                 yr, out, t, T, first, lds_d);
        first = load_path(pf_round_args(it, r + 1), gmem(it.bv), nullptr, gmem(it.ev),
                          gmem(it.num), Yl, p0);
      } else {
        ls_final_work<true, 8>(a, gmem(it.bv), nullptr, gmem(it.ev), gmem(it.num), beta, Y, yin,
                               yr, out, t, T, first);
        if (t == 0) {
          if (!sum_tri<true>(out, T, 1, tri, &flag)) goto fail;
          finalize_col<true>(out, T, 1, 0, 1.0 / static_cast<double>(P), gmem(it.price1));
        }
      }
    }
  }
  __syncthreads();

  // 4. frozen-strategy forward and 5. adjoint backward
  {
    const Model m = pf_model(it, it.k);
    const CoefTable cf{tab};
    switch (it.F) {
#define KOO_CASE(F)                                                                            \
  case F:                                                                                      \
    for (int vb = t; vb < it.nvb; vb += T) {                                                   \
      __syncthreads();                                                                         \
      berm_fwd_block<F, true>(m, it.b, cf, gmem(it.ckpt), gmem(it.lamt), gmem(it.tau),         \
                              gmem(it.vout), gmem(it.pfwd), vb);                               \
    }                                                                                          \
    for (int vb = t; vb < it.nvb; vb += T) {                                                   \
      __syncthreads();                                                                         \
      berm_bwd_block<F, true>(m, gmem(it.vols), gmem(it.ckpt), gmem(it.lamt), gmem(it.tau),    \
                              gmem(it.vout), gmem(it.scratch), gmem(it.pbwd), vb);             \
    }                                                                                          \
    break;
      KOO_CASE(1) KOO_CASE(2) KOO_CASE(3) KOO_CASE(4) KOO_CASE(5)
#undef KOO_CASE
    }
  }
  __syncthreads();

  // 6. final sums: the frozen price by the team's block 0, Greek column c by
  // block c mod T (each once all its partials are in)
  {
    const double scale = 1.0 / static_cast<double>(P);
    if (t == 0) {
      if (!wait_column(gmem(it.pfwd), it.nvb, 1, 0, &flag)) goto fail;
      finalize_col<true>(gmem(it.pfwd), it.nvb, 1, 0, scale, gmem(it.price));
    }
    for (int c = t; c < 2 * it.N; c += T) {
      if (!wait_column(gmem(it.pbwd), it.nvb, 2 * it.N, c, &flag)) goto fail;
      finalize_col<true>(gmem(it.pbwd), it.nvb, 2 * it.N, c, scale, gmem(it.greeks));
    }
  }
  if (threadIdx.x == 0) ticks[blockIdx.x] = __builtin_amdgcn_s_memrealtime() - t_start;
  return;
fail:
  if (threadIdx.x == 0) {
    ticks[blockIdx.x] = __builtin_amdgcn_s_memrealtime() - t_start;
    err[0] = 1u;
    *gmem(it.price) = __builtin_nan("");
    *gmem(it.price1) = __builtin_nan("");
  }
}

// Arms a product's exchange words (one block per product): the LS partial
// slots of its team, the forward's price partials and the backward's
// partials.
__global__ void __launch_bounds__(256) k_portfolio_arm(const PfItem* __restrict__ items) {
  const PfItem& it = items[blockIdx.x];
  const size_t nx = 3 * static_cast<size_t>(kTriMax) * it.team;
  const size_t nf = it.nvb;
  const size_t nbw = static_cast<size_t>(it.nvb) * 2 * it.N;
  unsigned long long* x = reinterpret_cast<unsigned long long*>(gmem(it.xbuf));
  unsigned long long* pf = reinterpret_cast<unsigned long long*>(gmem(it.pfwd));
  unsigned long long* pb = reinterpret_cast<unsigned long long*>(gmem(it.pbwd));
  for (size_t i = threadIdx.x; i < nx; i += blockDim.x) x[i] = kUnsetBits;
  for (size_t i = threadIdx.x; i < nf; i += blockDim.x) pf[i] = kUnsetBits;
  for (size_t i = threadIdx.x; i < nbw; i += blockDim.x) pb[i] = kUnsetBits;
}

// Device buffer: the items, the bridge tables, the forward price partials
// (one column per product, nvb doubles each), the error word.
constexpr int kPfMaxGrid = 4096;

struct PfLayout {
  size_t items, tables, pfwd, ticks, err, total;
};

PfLayout pf_layout(int n, size_t pfwd_doubles) {
  PfLayout l;
  auto al = [](size_t x) { return (x + 255) & ~static_cast<size_t>(255); };
  l.items = 0;
  l.tables = al(sizeof(PfItem) * n);
  l.pfwd = l.tables + al(sizeof(BridgeTables) * n);
  l.ticks = l.pfwd + al(sizeof(double) * pfwd_doubles);
  l.err = l.ticks + al(sizeof(unsigned long long) * kPfMaxGrid);
  l.total = l.err + 256;
  return l;
}

int pf_nvb(int paths) { return (paths + kPfThreads - 1) / kPfThreads; }

// Dynamic LDS (bytes) of the phases of a product: generation (S bridge
// points per lane and the S dimensions' direction numbers), path quantities
// and forward (the curve and the step table), backward (plus the vega slots),
// the LS rounds (kPfLsDoubles).
size_t pf_phase_lds(int N) {
  const size_t gen = sizeof(float) * static_cast<size_t>(N) * kPfThreads + sizeof(uint32_t) * 32 * N;
  size_t m = gen;
  m = std::max(m, fwd_lds(N, kPfThreads));
  m = std::max(m, bwd_lds(N, kPfThreads, true));
  m = std::max(m, sizeof(double) * static_cast<size_t>(kPfLsDoubles));
  return (m + 15) & ~static_cast<size_t>(15);
}

// Work of a product in rate updates per path, weighted by the per-update
// cost of each phase (generation per variate, evolution to the last exercise
// date in the path-quantity and forward kernels, the reverse sweep with its
// recompute, the LS rounds): the team sizes are proportional to it.
double pf_cost(const koo_berm_product& q, int last, int gen) {
  double u = 0.0;
  for (int s = 0; s <= last; ++s) u += q.rates - s;
  const double F = q.factors;
  const double evolve = u * (13.0 * F + 20.0);
  const double reverse = u * (34.0 * F + 60.0) + (q.checkpoint_every > 1 ? evolve : 0.0);
  const double g = gen ? 60.0 * q.rates * F : 0.0;
  return static_cast<double>(q.paths) * (2.0 * evolve + reverse + g + 400.0 * q.n_ex);
}

}  // namespace
}  // namespace koo

using namespace koo;

extern "C" size_t bermudan_portfolio_persistent_bytes(int n_products, const int* paths) {
  if (n_products < 1 || !paths) return 0;
  size_t f = 0;
  for (int i = 0; i < n_products; ++i) f += pf_nvb(paths[i] < 1 ? 1 : paths[i]);
  return pf_layout(n_products, f).total;
}

extern "C" int bermudan_portfolio_persistent_prepare(int n_products, const koo_gen_product* gen,
                                                     const koo_ls_product* ls,
                                                     const koo_berm_product* adj,
                                                     const double* work, void* device_buffer,
                                                     koo_persistent_plan* plan) {
  const char* who = "bermudan_portfolio_persistent_prepare";
  constexpr int kMaxProducts = sizeof(plan->team_first) / sizeof(plan->team_first[0]);
  if (n_products < 1 || n_products > kMaxProducts || !ls || !adj || !device_buffer || !plan)
    return set_error(KOO_ERR_INVALID, "%s: bad arguments (1 to %d products)", who, kMaxProducts);
  std::vector<PfItem> items(n_products);
  std::vector<BridgeTables> tables(n_products);
  std::vector<double> cost(n_products);
  size_t lds_main = 0, pf_total = 0;
  int max_e = 1;
  for (int i = 0; i < n_products; ++i) {
    const koo_ls_product& l = ls[i];
    const koo_berm_product& q = adj[i];
    if (q.paths < 1 || q.rates < 1 || q.rates > 45 || q.factors < 1 || q.factors > 5 ||
        q.n_ex < 1 || q.n_ex > kPfMaxDates || q.n_ex > q.rates || q.checkpoint_every < 1)
      return set_error(KOO_ERR_INVALID,
                       "%s: product %d: need paths>=1, 1<=rates<=45, 1<=factors<=5, "
                       "1<=dates<=min(rates,%d), checkpoint_every>=1",
                       who, i, kPfMaxDates);
    if (l.paths != q.paths || l.rates != q.rates || l.factors != q.factors || l.n_ex != q.n_ex ||
        l.variates != q.variates || l.initial_rates != q.initial_rates || l.taus != q.taus ||
        l.displacements != q.displacements || l.pseudo_roots != q.pseudo_roots ||
        l.strike != q.strike || l.itm_only != q.itm_only || l.coefficients != q.coefficients)
      return set_error(KOO_ERR_INVALID,
                       "%s: product %d: the LS and adjoint descriptions disagree (paths, rates, "
                       "factors, dates, model arrays, strike, itm_only and coefficients must be "
                       "the same)", who, i);
    if (!q.variates || !q.initial_rates || !q.taus || !q.displacements || !q.pseudo_roots ||
        !q.vols || !q.exercise_steps || !l.exercise_steps || !q.coefficients ||
        !q.checkpoints || !q.path_buffers || !q.segment_scratch || !q.partials ||
        !q.price_out || !q.greeks_out || !l.basis_variables || !l.exercise_values ||
        !l.numeraires || !l.workspace || !l.price_out)
      return set_error(KOO_ERR_INVALID, "%s: product %d: null pointer", who, i);
    for (int e = 0; e < q.n_ex; ++e)
      if (l.exercise_steps[e] != q.exercise_steps[e])
        return set_error(KOO_ERR_INVALID, "%s: product %d: exercise steps disagree", who, i);
    PfItem& it = items[i];
    it = PfItem{};
    if (gen) {
      const koo_gen_product& g = gen[i];
      if (!g.directions || !g.allocation_indices || (g.number_scrambled > 0 && !g.scramblers))
        return set_error(KOO_ERR_INVALID, "%s: product %d: null generator pointer", who, i);
      if (static_cast<uint64_t>(g.path_offset) + 1u + static_cast<uint64_t>(q.paths) > (1ull << 32))
        return set_error(KOO_ERR_INVALID, "%s: product %d: path_offset + paths exceeds 2^32 - 1",
                         who, i);
      it.gen = 1;
      it.n_scr = g.number_scrambled;
      it.off1 = g.path_offset + 1u;
      it.dirs = (const uint32_t KOO_GLOBAL*)g.directions;
      it.scr = (const uint32_t KOO_GLOBAL*)g.scramblers;
      it.alloc = (const int KOO_GLOBAL*)g.allocation_indices;
      tables[i] = tables_for(q.rates);
    }
    it.z = (g_float*)q.variates;
    it.f0 = (g_cfloat*)q.initial_rates;
    it.taus = (g_cfloat*)q.taus;
    it.disp = (g_cfloat*)q.displacements;
    it.A = (g_cfloat*)q.pseudo_roots;
    it.vols = (g_cfloat*)q.vols;
    it.strike = q.strike;
    it.P = q.paths;
    it.N = q.rates;
    it.F = q.factors;
    it.k = q.checkpoint_every;
    for (int s = 0; s < kMaxRates; ++s) it.ex.e_of_step[s] = -1;
    for (int e = 0; e < q.n_ex; ++e) {
      const int s = l.exercise_steps[e];
      if (s < 0 || s >= q.rates || (e > 0 && s <= l.exercise_steps[e - 1]))
        return set_error(KOO_ERR_INVALID, "%s: product %d: exercise steps must increase", who, i);
      it.ex.e_of_step[s] = static_cast<signed char>(e);
      for (int v = 0; v < 2; ++v) {
        it.sh[e][v] = l.var_shift ? l.var_shift[2 * e + v] : 0.0f;
        it.sc[e][v] = l.var_scale ? l.var_scale[2 * e + v] : 1.0f;
      }
    }
    it.ex.last = l.exercise_steps[q.n_ex - 1];
    it.E = q.n_ex;
    it.itm_only = q.itm_only != 0;
    it.bv = (g_float*)l.basis_variables;
    it.ev = (g_float*)l.exercise_values;
    it.num = (g_float*)l.numeraires;
    char* ws = static_cast<char*>(l.workspace);
    it.Y = (g_float*)ws;  // the running values lead the LS workspace
    it.xbuf = (g_double*)(ws + ls_workspace_xbuf_offset(q.paths));
    it.coeffs = (g_double*)q.coefficients;
    it.price1 = (g_double*)l.price_out;
    if (int e = make_berm_args(who, q.rates, q.n_ex, q.exercise_steps, q.var_shift, q.var_scale,
                               q.itm_only, &it.b))
      return e;
    it.ckpt = (g_float*)q.checkpoints;
    it.lamt = (g_float*)q.path_buffers;
    it.tau = (g_int*)(q.path_buffers + static_cast<size_t>(q.rates) * q.paths);
    it.vout = (g_float*)(q.path_buffers + static_cast<size_t>(q.rates + 1) * q.paths);
    it.scratch = (g_float*)q.segment_scratch;
    it.pbwd = (g_double*)q.partials;
    it.price = (g_double*)q.price_out;
    it.greeks = (g_double*)q.greeks_out;
    it.nvb = pf_nvb(q.paths);
    pf_total += it.nvb;
    lds_main = std::max(lds_main, pf_phase_lds(q.rates));
    max_e = std::max(max_e, q.n_ex);
    cost[i] = work ? work[i] : pf_cost(q, it.ex.last, it.gen);
    if (!(cost[i] > 0.0))
      return set_error(KOO_ERR_INVALID, "%s: product %d: work must be positive", who, i);
  }
  // dynamic LDS: the phases, then the FP32 coefficient table
  const size_t lds = lds_main + sizeof(float) * 6 * max_e;
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_portfolio, kPfThreads, lds) !=
          hipSuccess ||
      per_cu < 1)
    return set_error(KOO_ERR_INVALID, "%s: no block of %zu bytes of LDS can be resident", who, lds);
  // every block must be resident at once; bound the occupancy figure by the
  // residency measured on gfx950 for 256-lane blocks (LDS per block incl.
  // static: <= 40 KiB: 4 per CU, <= 53 120 B: 3, <= 80 KiB: 2), since the
  // query has allowed a block more than fits just below the 3-per-CU limit
  {
    hipFuncAttributes fa{};
    size_t total = lds;
    if (hipFuncGetAttributes(&fa, reinterpret_cast<const void*>(k_portfolio)) == hipSuccess)
      total += fa.sharedSizeBytes;
    const int by_lds = total <= 40960 ? 4 : (total <= 53120 ? 3 : (total <= 81920 ? 2 : 1));
    if (by_lds < per_cu) per_cu = by_lds;
  }
  // KOO_PF_BLOCKS_PER_CU caps the blocks per CU below the occupancy figure
  // (all blocks must be resident at once: the teams' waits rely on it)
  if (const char* cap_env = getenv("KOO_PF_BLOCKS_PER_CU")) {
    const int c = atoi(cap_env);
    if (c >= 1 && c < per_cu) per_cu = c;
  }
  const int R = std::min(per_cu * device_cu_count(), kPfMaxGrid);
  if (n_products > R)
    return set_error(KOO_ERR_INVALID, "%s: %d products exceed the %d resident blocks", who,
                     n_products, R);
  // team sizes: proportional to the work, at least 1, at most one block per
  // 256-path chunk and per LS partial column, summing to at most R
  double tot = 0.0;
  for (double c : cost) tot += c;
  std::vector<int> team(n_products);
  std::vector<int> cap(n_products);
  int used = 0;
  for (int i = 0; i < n_products; ++i) {
    cap[i] = std::min(items[i].nvb, ls_workspace_xbuf_columns(items[i].P));
    team[i] = std::max(1, std::min(cap[i], static_cast<int>(R * cost[i] / tot)));
    used += team[i];
  }
  while (used > R) {  // the floor of one block per product can overshoot
    int j = static_cast<int>(std::max_element(team.begin(), team.end()) - team.begin());
    --team[j];
    --used;
  }
  for (bool grew = true; grew && used < R;) {  // spare blocks to the most loaded per block
    grew = false;
    int j = -1;
    double worst = 0.0;
    for (int i = 0; i < n_products; ++i)
      if (team[i] < cap[i] && cost[i] / team[i] > worst) {
        worst = cost[i] / team[i];
        j = i;
      }
    if (j >= 0) {
      ++team[j];
      ++used;
      grew = true;
    }
  }
  const PfLayout lay = pf_layout(n_products, pf_total);
  char* buf = static_cast<char*>(device_buffer);
  size_t fo = 0;
  int b0 = 0;
  for (int i = 0; i < n_products; ++i) {
    PfItem& it = items[i];
    it.team0 = b0;
    it.team = team[i];
    b0 += team[i];
    it.pfwd = (g_double*)(buf + lay.pfwd) + fo;
    fo += it.nvb;
    if (it.gen)
      it.tables = (const BridgeTables KOO_GLOBAL*)(buf + lay.tables) + i;
  }
  if (hipMemcpy(buf + lay.items, items.data(), sizeof(PfItem) * n_products,
                hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(buf + lay.tables, tables.data(), sizeof(BridgeTables) * n_products,
                hipMemcpyHostToDevice) != hipSuccess)
    return set_error(KOO_ERR_HIP, "%s: descriptor copy failed", who);
  *plan = koo_persistent_plan{};
  plan->products = n_products;
  plan->grid = b0;
  plan->lds_bytes = static_cast<int>(lds);
  plan->tab_offset = static_cast<int>(lds_main / sizeof(float));
  plan->blocks_per_cu = per_cu;
  plan->err_offset = lay.err;
  plan->ticks_offset = lay.ticks;
  for (int i = 0; i < n_products; ++i) {
    plan->team_first[i] = items[i].team0;
    plan->team_blocks[i] = items[i].team;
  }
  return KOO_OK;
}

extern "C" int bermudan_portfolio_persistent_run(const koo_persistent_plan* plan,
                                                 void* device_buffer, void* stream) {
  const char* who = "bermudan_portfolio_persistent_run";
  if (!plan || !device_buffer || plan->products < 1 || plan->grid < plan->products ||
      plan->lds_bytes < 1 || plan->tab_offset < 1)
    return set_error(KOO_ERR_INVALID, "%s: bad plan", who);
  char* buf = static_cast<char*>(device_buffer);
  const PfItem* items = reinterpret_cast<const PfItem*>(buf);
  // the error word sits after the items, tables and forward partials; its
  // offset is recomputed from the item table on the device side's layout
  unsigned* err = reinterpret_cast<unsigned*>(buf + plan->err_offset);
  hipStream_t st = as_stream(stream);
  hipLaunchKernelGGL(k_portfolio_arm, dim3(plan->products), dim3(256), 0, st, items);
  hipLaunchKernelGGL(k_zero_words, dim3(1), dim3(64), 0, st, err, 1);
  unsigned long long* ticks = reinterpret_cast<unsigned long long*>(buf + plan->ticks_offset);
  hipLaunchKernelGGL(k_portfolio, dim3(plan->grid), dim3(kPfThreads), plan->lds_bytes, st, items,
                     plan->products, plan->tab_offset, err, ticks);
  return check_launch(who);
}
