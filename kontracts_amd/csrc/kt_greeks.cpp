// kt_greeks.cpp — Greeks by CRN bump-and-reprice (src/greeks.rs:60-98).
#include <cmath>
#include <limits>

#include "kt_engine_internal.h"

namespace kt {

namespace {

double price_once(kt_engine* E, const Con& con, const ModelDesc& md, const std::vector<double>& grid, double rate,
                  const kt_mc_config* cfg, double* se) {
  validate_model(md);
  Program prog = lower({&con}, grid, md.names);
  auto st = run_program(E, prog, md, grid, rate, cfg->n_paths, cfg->seed, nullptr, nullptr);
  kt_price_result r = result_from_stats(st[0]);
  if (se) *se = r.std_error;
  return r.price;
}

}  // namespace

// greeks_gbm: drift = rate, seven repricings at the same seed.  The seven
// scenarios run in one launch that shares every normal (common random
// numbers); the kernel reduces the per-path samples of price, delta, gamma,
// vega and rho, whose means are exactly the reference's finite differences of
// the seven prices, and whose spread gives each Greek's standard error.
void greeks_gbm_crn(kt_engine* E, const char* json, const char* asset, double s0, double sigma,
                    const kt_mc_config* cfg, const kt_bumps& h, kt_greeks* out) {
  if (cfg->antithetic || cfg->control_variate)
    raise(KT_ERR_UNSUPPORTED, "variance reduction is not on the GPU engine's path");
  auto con = parse_contract(json);
  Plan plan = compile_plan(*con);
  std::vector<double> grid = cfg->grid && cfg->n_grid ? std::vector<double>(cfg->grid, cfg->grid + cfg->n_grid)
                                                      : time_grid_forced(plan, cfg->steps_per_year, false);
  auto model = [&](double s, double sg, double r) {
    ModelDesc md;
    md.kind = KT_MODEL_GBM;
    md.names = {asset};
    md.s0 = {s};
    md.mu = {r};
    md.sigma = {sg};
    return md;
  };
  const double r = cfg->rate;
  std::vector<ModelDesc> scen = {model(s0, sigma, r),         model(s0 + h.spot, sigma, r),
                                 model(s0 - h.spot, sigma, r), model(s0, sigma + h.vol, r),
                                 model(s0, sigma - h.vol, r),  model(s0, sigma, r + h.rate),
                                 model(s0, sigma, r - h.rate)};
  std::vector<double> rates = {r, r, r, r, r, r + h.rate, r - h.rate};
  for (const ModelDesc& m : scen) validate_model(m);
  validate_grid(grid);
  *out = kt_greeks{};
  if (cfg->n_paths == 0) return;
  Program prog = lower({con.get()}, grid, {asset});
  int ns = 0, nout = 0;
  if (path_variant_exists(MK_GBM, 1, prog.n_slots, 1, false, kScenarios, &ns, &nout)) {
    const double hs = h.spot, hv = h.vol, hr = h.rate;
    std::vector<std::vector<double>> w = {
        {1, 0, 0, 0, 0, 0, 0},
        {0, 1 / (2 * hs), -1 / (2 * hs), 0, 0, 0, 0},
        {-2 / (hs * hs), 1 / (hs * hs), 1 / (hs * hs), 0, 0, 0, 0},
        {0, 0, 0, 1 / (2 * hv), -1 / (2 * hv), 0, 0},
        {0, 0, 0, 0, 0, 1 / (2 * hr), -1 / (2 * hr)}};
    auto st = run_program_scen(E, prog, scen, rates, w, grid, cfg->n_paths, cfg->seed, nullptr, nullptr);
    double* vals[5] = {&out->price, &out->delta, &out->gamma, &out->vega, &out->rho};
    for (int j = 0; j < 5; ++j) {
      kt_price_result pr = result_from_stats(st[j]);
      *vals[j] = pr.price;
      out->std_error[j] = pr.std_error;
    }
    return;
  }
  // Programs too large for the scenario kernel: seven launches at the same seed.
  double p[7];
  double se0 = 0.0;
  for (int i = 0; i < 7; ++i) {
    auto st = run_program(E, prog, scen[i], grid, rates[i], cfg->n_paths, cfg->seed, nullptr, nullptr);
    kt_price_result pr = result_from_stats(st[0]);
    p[i] = pr.price;
    if (i == 0) se0 = pr.std_error;
  }
  out->price = p[0];
  out->delta = (p[1] - p[2]) / (2.0 * h.spot);
  out->gamma = (p[1] - 2.0 * p[0] + p[2]) / (h.spot * h.spot);
  out->vega = (p[3] - p[4]) / (2.0 * h.vol);
  out->rho = (p[5] - p[6]) / (2.0 * h.rate);
  const double nan = std::numeric_limits<double>::quiet_NaN();
  out->std_error[0] = se0;
  for (int i = 1; i < 5; ++i) out->std_error[i] = nan;
}

// Central differences over the model parameters with common random numbers.
void greeks_bump(kt_engine* E, const char* json, const kt_model* model, const kt_mc_config* cfg, const double* h,
                 uint64_t n_params, double* out, double* stderr_out) {
  ModelDesc base = model_from_abi(model);
  PreparedContract pc = prepare(json, base, cfg);
  uint64_t want = base.kind == KT_MODEL_HESTON ? 6 : (base.kind == KT_MODEL_GBM ? 3 : 0);
  if (want == 0) raise(KT_ERR_UNSUPPORTED, "bump Greeks are defined for GBM and Heston models");
  if (n_params != want) raise(KT_ERR_INVALID_ARGUMENT, "wrong number of bump sizes for this model");
  for (uint64_t i = 0; i <= n_params; ++i) out[i] = 0.0;
  if (cfg->n_paths == 0) return;
  auto bumped = [&](int which, double delta) {
    ModelDesc m = base;
    if (base.kind == KT_MODEL_GBM) {
      if (which == 0) m.s0[0] += delta;
      if (which == 1) m.sigma[0] += delta;
      if (which == 2) m.mu[0] += delta;
    } else {
      if (which == 0) m.s0[0] += delta;
      if (which == 1) m.v0 += delta;
      if (which == 2) m.kappa += delta;
      if (which == 3) m.theta += delta;
      if (which == 4) m.sigma_v += delta;
      if (which == 5) m.rho += delta;
    }
    return m;
  };
  double se0 = 0.0;
  out[0] = price_once(E, *pc.con, base, pc.grid, cfg->rate, cfg, &se0);
  for (uint64_t j = 0; j < n_params; ++j) {
    double up = price_once(E, *pc.con, bumped((int)j, h[j]), pc.grid, cfg->rate, cfg, nullptr);
    double dn = price_once(E, *pc.con, bumped((int)j, -h[j]), pc.grid, cfg->rate, cfg, nullptr);
    out[1 + j] = (up - dn) / (2.0 * h[j]);
  }
  if (stderr_out) {
    stderr_out[0] = se0;
    for (uint64_t j = 0; j < n_params; ++j) stderr_out[1 + j] = std::numeric_limits<double>::quiet_NaN();
  }
}

void greeks_adjoint(kt_engine*, const char*, const kt_model*, const kt_mc_config*, double*, double*, uint64_t) {
  raise(KT_ERR_UNSUPPORTED, "pathwise Greeks: not built yet");
}

void price_portfolio(kt_engine*, const kt_job*, uint64_t, kt_price_result*) {
  raise(KT_ERR_UNSUPPORTED, "portfolio: not built yet");
}

}  // namespace kt
