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

// greeks_gbm: drift = rate, seven repricings at the same seed.
void greeks_gbm_crn(kt_engine* E, const char* json, const char* asset, double s0, double sigma,
                    const kt_mc_config* cfg, const kt_bumps& h, kt_greeks* out) {
  auto con = parse_contract(json);
  Plan plan = compile_plan(*con);
  std::vector<double> grid = cfg->grid && cfg->n_grid ? std::vector<double>(cfg->grid, cfg->grid + cfg->n_grid)
                                                      : time_grid_forced(plan, cfg->steps_per_year, false);
  validate_grid(grid);
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
  *out = kt_greeks{};
  if (cfg->n_paths == 0) return;
  double se0 = 0.0;
  double p = price_once(E, *con, model(s0, sigma, r), grid, r, cfg, &se0);
  double su = price_once(E, *con, model(s0 + h.spot, sigma, r), grid, r, cfg, nullptr);
  double sd = price_once(E, *con, model(s0 - h.spot, sigma, r), grid, r, cfg, nullptr);
  double vu = price_once(E, *con, model(s0, sigma + h.vol, r), grid, r, cfg, nullptr);
  double vd = price_once(E, *con, model(s0, sigma - h.vol, r), grid, r, cfg, nullptr);
  double ru = price_once(E, *con, model(s0, sigma, r + h.rate), grid, r + h.rate, cfg, nullptr);
  double rd = price_once(E, *con, model(s0, sigma, r - h.rate), grid, r - h.rate, cfg, nullptr);
  out->price = p;
  out->delta = (su - sd) / (2.0 * h.spot);
  out->gamma = (su - 2.0 * p + sd) / (h.spot * h.spot);
  out->vega = (vu - vd) / (2.0 * h.vol);
  out->rho = (ru - rd) / (2.0 * h.rate);
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

void price_american(kt_engine*, const char*, const double*, uint64_t, const kt_model*, const kt_mc_config*, int,
                    kt_price_result*) {
  raise(KT_ERR_UNSUPPORTED, "LSM: not built yet");
}

void price_portfolio(kt_engine*, const kt_job*, uint64_t, kt_price_result*) {
  raise(KT_ERR_UNSUPPORTED, "portfolio: not built yet");
}

}  // namespace kt
