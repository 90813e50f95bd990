// pt_opt.hip -- the temperature-set update of tamc-pt-opt (src/bin/tamc-pt-opt.rs:34-198) and the
// tamc-core helpers it uses (tamc-core/src/util.rs:6-138), in f32 with the reference's operation
// order (no contraction: the library is built with -ffp-contract=off).  Host code: the PT runs it
// consumes come from the batched kernel.
#include <math.h>
#include <string.h>

#include <string>
#include <vector>

#include "runtime_common.hpp"

using namespace tamc;

namespace {

// ndarray's sum of a contiguous array: eight partial sums, combined pairwise, then the tail
float unrolled_sum(const std::vector<float> &xs) {
  float p[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  size_t k = 0;
  for (; xs.size() - k >= 8; k += 8)
    for (int a = 0; a < 8; ++a) p[a] = p[a] + xs[k + a];
  float acc = 0.f;
  for (int a = 0; a < 4; ++a) acc = acc + (p[a] + p[a + 4]);
  for (; k < xs.size(); ++k) acc = acc + xs[k];
  return acc;
}

// StepwiseMeasure::new (tamc-core/src/util.rs:24-43)
std::vector<float> stepwise_weights(const std::vector<float> &x) {
  const size_t n = x.size();
  std::vector<float> dx(n - 1), w(n, 0.f);
  for (size_t i = 0; i + 1 < n; ++i) dx[i] = x[i + 1] - x[i];
  w[0] = dx[0] / 2.0f;
  for (size_t i = 1; i + 1 < n; ++i) w[i] = (dx[i - 1] + dx[i]) / 2.0f;
  w[n - 1] = dx[n - 2] / 2.0f;
  return w;
}

// finite_differences (tamc-core/src/util.rs:6-22)
std::vector<float> finite_differences(const std::vector<float> &x, const std::vector<float> &y) {
  const size_t n = x.size();
  std::vector<float> dx(n - 1), dy(n - 1), d(n, 0.f);
  for (size_t i = 0; i + 1 < n; ++i) {
    dx[i] = x[i + 1] - x[i];
    dy[i] = y[i + 1] - y[i];
  }
  d[0] = dy[0] / dx[0];
  d[n - 1] = dy[n - 2] / dx[n - 2];
  for (size_t i = 1; i + 1 < n; ++i) d[i] = (dy[i - 1] * dx[i] / dx[i - 1] + dy[i] * dx[i - 1] / dx[i]) / (dx[i - 1] + dx[i]);
  return d;
}

// piecewise-linear interpolation, clamped at the ends (the `interp` crate's function)
float interp(const std::vector<float> &x, const std::vector<float> &y, float xp) {
  if (xp <= x.front()) return y.front();
  if (xp >= x.back()) return y.back();
  size_t i = 0;
  while (i + 2 < x.size() && x[i + 1] <= xp) ++i;
  return y[i] + (xp - x[i]) * (y[i + 1] - y[i]) / (x[i + 1] - x[i]);
}

struct Eta {
  const std::vector<float> &beta, &cdf;
  float operator()(float x) const {
    const float bmin = beta.front(), bmax = beta.back();
    return interp(beta, cdf, bmin + x * (bmax - bmin));
  }
};

// monotonic_bisection (tamc-core/src/util.rs:62-92): 0 root, 1 lower bound, 2 upper bound, 3 failed
int monotonic_bisection(const Eta &f, float y0, float xmin, float xmax, float *root) {
  const float tol = 1.0e-6f;
  float x0 = xmin, x1 = xmax;
  for (uint32_t it = 0; it < 1000000u; ++it) {
    const float xmid = (x0 + x1) / 2.0f;
    const float y = f(xmid);
    if (fabsf(y - y0) < tol) {
      *root = xmid;
      return 0;
    }
    if (fabsf(xmid - xmin) < tol) return 1;
    if (fabsf(xmid - xmax) < tol) return 2;
    if (y < y0) x0 = xmid;
    else x1 = xmid;
  }
  return 3;
}

// bisect_array (tamc-core/src/util.rs:96-125)
bool bisect_array(const Eta &f, float *roots, const float *ytgt, size_t n, float xmin, float xmax, std::string &err) {
  if (n == 0) return true;
  if (n == 1) {
    float r = 0.f;
    const int k = monotonic_bisection(f, ytgt[0], xmin, xmax, &r);
    if (k != 0) {
      static const char *what[] = {"", "terminated at the lower bound", "terminated at the upper bound", "failed"};
      err = "Bisection for " + std::to_string(ytgt[0]) + " in the interval [" + std::to_string(xmin) + "," +
            std::to_string(xmax) + "] " + what[k];
      return false;
    }
    roots[0] = r;
    return true;
  }
  const size_t mid = n / 2;
  if (!bisect_array(f, roots + mid, ytgt + mid, 1, xmin, xmax, err)) return false;
  const float xmid = roots[mid];
  return bisect_array(f, roots, ytgt, mid, xmin, xmid, err) &&
         bisect_array(f, roots + mid + 1, ytgt + mid + 1, n - mid - 1, xmid, xmax, err);
}

}  // namespace

extern "C" {

tamc_status tamc_pt_opt_update(uint32_t nt, const float *beta_vec, float *momentum, uint32_t count,
                               const uint32_t *dif_hists, const uint32_t *round_trips, uint32_t num_sweeps,
                               uint32_t num_chains, float alpha, float m, float *new_betas, float *err_out,
                               float *taus, int *updated) {
  clear_error();
  if (!beta_vec || !momentum || !new_betas || !err_out || !updated || (count && (!dif_hists || !round_trips || !taus)))
    return fail(TAMC_E_INVALID_ARG, "null argument");
  if (nt < 3) return fail(TAMC_E_INVALID_ARG, "the temperature optimisation needs at least three betas");
  if (count == 0) return fail(TAMC_E_INVALID_ARG, "the temperature optimisation needs at least one instance");
  const std::vector<float> beta(beta_vec, beta_vec + nt);
  const std::vector<float> weights = stepwise_weights(beta);
  std::vector<float> difs(nt - 1);
  for (uint32_t i = 0; i + 1 < nt; ++i) difs[i] = beta[i + 1] - beta[i];
  std::vector<float> tau(count), w(nt, 0.f);
  std::vector<std::vector<float>> ddif(count);
  for (uint32_t k = 0; k < count; ++k) {
    const uint32_t rts = round_trips[k];
    const float rt_per_rep = (float)rts / (float)(num_chains * nt);
    tau[k] = rts == 0 ? 0.0f : (float)num_sweeps / rt_per_rep;
    std::vector<float> probs(nt);
    for (uint32_t b = 0; b < nt; ++b) {
      const float a0 = (float)(dif_hists[((size_t)k * nt + b) * 2] + 1u);
      const float a1 = (float)(dif_hists[((size_t)k * nt + b) * 2 + 1] + 1u);
      probs[b] = a1 / (a0 + a1);
    }
    ddif[k] = finite_differences(beta, probs);
  }
  const float sum_tau = unrolled_sum(tau);
  for (uint32_t k = 0; k < count; ++k)
    for (uint32_t i = 0; i < nt; ++i) w[i] = w[i] + tau[k] * ddif[k][i];
  for (uint32_t i = 0; i < nt; ++i) w[i] = w[i] / sum_tau;
  std::vector<float> eta(nt), eta_w(nt - 1);
  for (uint32_t i = 0; i < nt; ++i) eta[i] = sqrtf(fmaxf(w[i] / weights[i], 0.0f));
  for (uint32_t i = 0; i + 1 < nt; ++i) eta_w[i] = (eta[i] + eta[i + 1]) / 2.0f * difs[i];
  for (uint32_t k = 0; k < count; ++k) taus[k] = tau[k];
  const float z = unrolled_sum(eta_w);
  if (z < 1.1920929e-07f) {  // f32::EPSILON: the reference warns and keeps the current set
    *updated = 0;
    memcpy(new_betas, beta_vec, nt * sizeof(float));
    *err_out = 0.f;
    return TAMC_OK;
  }
  std::vector<float> cdf;
  float acc = 0.f;
  for (float x : eta_w) {
    cdf.push_back(acc);
    acc = acc + x / z;
  }
  cdf.push_back(1.0f);
  // monotonic_divisions (tamc-core/src/util.rs:127-138) of eta's CDF into nt - 1 equal steps
  const uint32_t mdiv = nt - 1;
  std::vector<float> parts(mdiv - 1), xd(nt, 0.f);
  for (uint32_t i = 1; i < mdiv; ++i) parts[i - 1] = (float)((double)i / (double)mdiv);
  xd[mdiv] = 1.0f;
  std::string err;
  const Eta f{beta, cdf};
  if (!bisect_array(f, xd.data() + 1, parts.data(), mdiv - 1, 0.0f, 1.0f, err)) return fail(TAMC_E_INVALID_ARG, err);
  const float bmin = beta.front(), bmax = beta.back();
  std::vector<float> calc(nt);
  for (uint32_t i = 0; i < nt; ++i) calc[i] = xd[i] * (bmax - bmin) + bmin;
  for (uint32_t i = 0; i < nt; ++i) momentum[i] = expf(m * logf(momentum[i]) + (1.0f - m) * logf(calc[i]));
  float e = 0.f;
  for (uint32_t i = 0; i < nt; ++i) new_betas[i] = calc[i];
  for (uint32_t i = 1; i + 1 < nt; ++i) {
    const float b = expf(alpha * logf(momentum[i]) + (1.0f - alpha) * logf(beta[i]));
    e += fabsf(log10f(calc[i]) - log10f(beta[i]));
    new_betas[i] = b;
  }
  *err_out = e / (float)nt;
  *updated = 1;
  return TAMC_OK;
}

}  // extern "C"
