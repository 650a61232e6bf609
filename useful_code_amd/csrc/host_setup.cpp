// host_setup.cpp -- host-side model set-up for the LMM pricing path: the
// correlation model, PCA pseudo-roots per evolution step, fixed and spot
// drifts.  FP64 throughout; the device kernels take FP32 copies.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/koo_c_abi.h"
#include "koo_host.h"

namespace {

// Symmetric eigen-decomposition by cyclic Jacobi rotations with a shrinking
// threshold; eigenvalues descending, each eigenvector signed so that its
// largest-magnitude component is positive (a deterministic PCA sign).
void jacobi_eigen(int n, std::vector<double> a, std::vector<double>& evals,
                  std::vector<double>& evecs) {
  std::vector<double> v(static_cast<size_t>(n) * n, 0.0);
  for (int i = 0; i < n; ++i) v[static_cast<size_t>(i) * n + i] = 1.0;
  auto A = [&](int i, int j) -> double& { return a[static_cast<size_t>(i) * n + j]; };
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, diag = 0.0;
    for (int i = 0; i < n; ++i) {
      diag += A(i, i) * A(i, i);
      for (int j = i + 1; j < n; ++j) off += A(i, j) * A(i, j);
    }
    if (off <= 1e-32 * std::max(diag, 1e-300)) break;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        const double apq = A(p, q);
        if (std::fabs(apq) < 1e-300) continue;
        const double theta = (A(q, q) - A(p, p)) / (2.0 * apq);
        const double t = (theta >= 0.0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; ++k) {  // columns p, q
          const double akp = A(k, p), akq = A(k, q);
          A(k, p) = c * akp - s * akq;
          A(k, q) = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {  // rows p, q
          const double apk = A(p, k), aqk = A(q, k);
          A(p, k) = c * apk - s * aqk;
          A(q, k) = s * apk + c * aqk;
        }
        for (int k = 0; k < n; ++k) {
          const double vkp = v[static_cast<size_t>(k) * n + p], vkq = v[static_cast<size_t>(k) * n + q];
          v[static_cast<size_t>(k) * n + p] = c * vkp - s * vkq;
          v[static_cast<size_t>(k) * n + q] = s * vkp + c * vkq;
        }
      }
  }
  std::vector<int> order(n);
  for (int i = 0; i < n; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return A(x, x) > A(y, y); });
  evals.assign(n, 0.0);
  evecs.assign(static_cast<size_t>(n) * n, 0.0);
  for (int c = 0; c < n; ++c) {
    const int src = order[c];
    evals[c] = A(src, src);
    int imax = 0;
    for (int k = 1; k < n; ++k)
      if (std::fabs(v[static_cast<size_t>(k) * n + src]) >
          std::fabs(v[static_cast<size_t>(imax) * n + src]) + 1e-12)
        imax = k;
    const double sg = v[static_cast<size_t>(imax) * n + src] < 0.0 ? -1.0 : 1.0;
    for (int k = 0; k < n; ++k) evecs[static_cast<size_t>(k) * n + c] = sg * v[static_cast<size_t>(k) * n + src];
  }
}

}  // namespace

namespace koo {

void exp_corr(const double* times, int n, double L, double beta, double* corr) {
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j)
      corr[static_cast<size_t>(i) * n + j] =
          i == j ? 1.0 : L + (1.0 - L) * std::exp(-beta * std::fabs(times[i] - times[j]));
}

void pseudo_sqrt(int n, const double* cov, int factors, bool normalise, double* out) {
  std::vector<double> evals, evecs;
  jacobi_eigen(n, std::vector<double>(cov, cov + static_cast<size_t>(n) * n), evals, evecs);
  for (int f = 0; f < factors; ++f) {
    const double rl = evals[f] > 0.0 ? std::sqrt(evals[f]) : 0.0;
    for (int i = 0; i < n; ++i) out[static_cast<size_t>(i) * factors + f] = evecs[static_cast<size_t>(i) * n + f] * rl;
  }
  if (!normalise) return;
  for (int i = 0; i < n; ++i) {
    double nrm = 0.0;
    for (int f = 0; f < factors; ++f) nrm += out[static_cast<size_t>(i) * factors + f] * out[static_cast<size_t>(i) * factors + f];
    if (nrm > 0.0) {
      const double adj = std::sqrt(cov[static_cast<size_t>(i) * n + i] / nrm);
      for (int f = 0; f < factors; ++f) out[static_cast<size_t>(i) * factors + f] *= adj;
    }
  }
}

}  // namespace koo

using namespace koo;

extern "C" int koo_exp_corr(const double* times, int n, double L, double beta, double* corr) {
  if (n < 1 || !times || !corr) return set_host_error(KOO_ERR_INVALID, "koo_exp_corr: bad arguments");
  exp_corr(times, n, L, beta, corr);
  return KOO_OK;
}

extern "C" int koo_cholesky_lower(int n, const double* corr, double* L) {
  if (n < 1 || !corr || !L) return set_host_error(KOO_ERR_INVALID, "koo_cholesky_lower: bad arguments");
  for (int i = 0; i < n * n; ++i) L[i] = 0.0;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j <= i; ++j) {
      double s = corr[i * n + j];
      for (int k = 0; k < j; ++k) s -= L[i * n + k] * L[j * n + k];
      if (i == j) {
        if (!(s > 0.0))
          return set_host_error(KOO_ERR_INVALID,
                                "koo_cholesky_lower: matrix not positive definite at row %d", i);
        L[i * n + i] = std::sqrt(s);
      } else {
        L[i * n + j] = s / L[j * n + j];
      }
    }
  return KOO_OK;
}

extern "C" int koo_pseudo_sqrt(int n, const double* cov, int factors, int normalise, double* out) {
  if (n < 1 || factors < 1 || factors > n || !cov || !out)
    return set_host_error(KOO_ERR_INVALID, "koo_pseudo_sqrt: bad arguments");
  pseudo_sqrt(n, cov, factors, normalise != 0, out);
  return KOO_OK;
}

// Per-step PCA pseudo-roots of the flat-vol covariance accrued over each
// evolution step by the rates not yet reset; the construction of
// gold/volstructs_gold.cpp:177-258 (FlatVolPseudoRootsOfCovariances with
// GenerateExpLongCorrMatrix on the rate start times).
extern "C" int FlatVolPseudoRootsOfCovariances(const double* rate_times, const double* evol_times,
                                               const double* vols, int n_rates, int steps,
                                               int factors, double L, double beta, double* A) {
  if (n_rates < 1 || steps < 1 || factors < 1 || factors > n_rates || !rate_times ||
      !evol_times || !vols || !A)
    return set_host_error(KOO_ERR_INVALID, "FlatVolPseudoRootsOfCovariances: bad arguments");
  const size_t nn = static_cast<size_t>(n_rates) * n_rates;
  std::vector<double> corr(nn), cov(nn);
  exp_corr(rate_times, n_rates, L, beta, corr.data());
  double start = 0.0;
  for (int s = 0; s < steps; ++s) {
    for (int r = 0; r < n_rates; ++r) {
      const double tr = std::min(evol_times[s], rate_times[r]) - std::min(start, rate_times[r]);
      for (int q = 0; q <= r; ++q) {
        const double tq = std::min(evol_times[s], rate_times[q]) - std::min(start, rate_times[q]);
        const double c = corr[static_cast<size_t>(r) * n_rates + q] * vols[r] * vols[q] * std::min(tr, tq);
        cov[static_cast<size_t>(r) * n_rates + q] = cov[static_cast<size_t>(q) * n_rates + r] = c;
      }
    }
    start = evol_times[s];
    pseudo_sqrt(n_rates, cov.data(), factors, true, A + static_cast<size_t>(s) * n_rates * factors);
  }
  return KOO_OK;
}

// A(s,r,f) = sigma_r sqrt(T_s - T_{s-1}) B(r,f) for r >= s (T_{-1} = 0),
// B the row-renormalised rank-F PCA root of the rate correlation; evolution
// times are the rate start times.  Vol-row scaling makes dA/dsigma_r = A/sigma_r.
extern "C" int koo_factor_pseudo_roots(const double* rate_times, const double* vols, int n_rates,
                                       int factors, double L, double beta, double* A, double* B) {
  if (n_rates < 1 || factors < 1 || factors > n_rates || !rate_times || !vols || !A)
    return set_host_error(KOO_ERR_INVALID, "koo_factor_pseudo_roots: bad arguments");
  const size_t nn = static_cast<size_t>(n_rates) * n_rates;
  std::vector<double> corr(nn), b(static_cast<size_t>(n_rates) * factors);
  exp_corr(rate_times, n_rates, L, beta, corr.data());
  pseudo_sqrt(n_rates, corr.data(), factors, true, b.data());
  double prev = 0.0;
  for (int s = 0; s < n_rates; ++s) {
    const double sq = std::sqrt(rate_times[s] - prev);
    prev = rate_times[s];
    for (int r = 0; r < n_rates; ++r)
      for (int f = 0; f < factors; ++f)
        A[(static_cast<size_t>(s) * n_rates + r) * factors + f] =
            r >= s ? vols[r] * sq * b[static_cast<size_t>(r) * factors + f] : 0.0;
  }
  if (B) std::memcpy(B, b.data(), sizeof(double) * b.size());
  return KOO_OK;
}

// Fixed drifts -0.5 sum_f A(s,r,f)^2  (gold/LMM_evolver_classes_gold.h:193-205).
extern "C" int koo_fixed_drifts(const double* A, int steps, int n_rates, int factors, double* out) {
  if (steps < 1 || n_rates < 1 || factors < 1 || !A || !out)
    return set_host_error(KOO_ERR_INVALID, "koo_fixed_drifts: bad arguments");
  for (int s = 0; s < steps; ++s)
    for (int r = 0; r < n_rates; ++r) {
      double v = 0.0;
      for (int f = 0; f < factors; ++f) {
        const double x = A[(static_cast<size_t>(s) * n_rates + r) * factors + f];
        v += x * x;
      }
      out[static_cast<size_t>(s) * n_rates + r] = -0.5 * v;
    }
  return KOO_OK;
}

// Spot-measure drifts of one step from the rates at its start, summed from
// the alive index; spotDriftComputer::getDrifts (gold/LMM_evolver_gold.h:
// 178-231; the alive variant there seeds with g[0], here with g[alive]).
extern "C" int spotDriftComputer_getDrifts(const double* A_step, const double* rates,
                                           const double* taus, const double* displacements,
                                           int factors, int n_rates, int alive, double* drifts) {
  if (n_rates < 1 || factors < 1 || alive < 0 || alive > n_rates || !A_step || !rates || !taus ||
      !displacements || !drifts)
    return set_host_error(KOO_ERR_INVALID, "spotDriftComputer_getDrifts: bad arguments");
  std::vector<double> e(factors, 0.0);
  for (int r = 0; r < alive; ++r) drifts[r] = 0.0;
  for (int r = alive; r < n_rates; ++r) {
    const double g = (rates[r] + displacements[r]) * taus[r] / (1.0 + rates[r] * taus[r]);
    double mu = 0.0;
    for (int f = 0; f < factors; ++f) {
      e[f] += A_step[static_cast<size_t>(r) * factors + f] * g;
      mu += A_step[static_cast<size_t>(r) * factors + f] * e[f];
    }
    drifts[r] = mu;
  }
  return KOO_OK;
}
