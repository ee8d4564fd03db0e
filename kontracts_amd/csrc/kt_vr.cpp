// kt_vr.cpp — variance reduction on the path kernels (src/pricer.rs:117-228,
// src/variance_reduction.rs:150-235).
//
// * Antithetic pairs (GBM only, like the reference's simulators): n_paths / 2
//   pairs, each lane of the path kernel simulating the path on its stream's
//   normals and the path on their negations, and keeping the pair mean.
// * Control variate: an at-the-money call on the model's first asset paying
//   at the grid's last date, priced on the same paths in the same launch as a
//   second output; beta = 1 and its Black-Scholes value corrects the price.
//   Without antithetic pairs the kernel reduces the per-path difference
//   pv_target - pv_control directly.
#include <cmath>
#include <cstdio>
#include <limits>
#include <string>

#include "kt_engine_internal.h"

namespace kt {

namespace {

// N(x) through the Abramowitz & Stegun 7.1.26 erf approximation (|error| <=
// 1.5e-7), the normal CDF behind the reference's Black-Scholes price
// (src/numerics.rs:22-41).  The control variate's correction is this price,
// so it is evaluated the same way, in the same operation order.
double norm_cdf_as(double x) {
  const double z = x / std::sqrt(2.0);
  const double sign = z < 0.0 ? -1.0 : 1.0;
  const double za = std::fabs(z);
  const double t = 1.0 / (1.0 + 0.3275911 * za);
  const double poly = ((((1.061405429 * t - 1.453152027) * t) + 1.421413741) * t - 0.284496736) * t + 0.254829592;
  const double erf_abs = 1.0 - poly * t * std::exp(-za * za);
  return 0.5 * (1.0 + sign * erf_abs);
}

// Black-Scholes call (src/numerics.rs:63-71).
double bs_call_as(double s, double k, double t, double r, double sigma) {
  if (t <= 0.0 || sigma <= 0.0) return std::max(s - k * std::exp(-r * t), 0.0);
  const double vst = sigma * std::sqrt(t);
  const double d1 = (std::log(s / k) + (r + 0.5 * sigma * sigma) * t) / vst;
  const double d2 = d1 - vst;
  return s * norm_cdf_as(d1) - k * std::exp(-r * t) * norm_cdf_as(d2);
}

std::string json_string(const std::string& s) {
  std::string o = "\"";
  for (unsigned char c : s) {
    if (c == '"' || c == '\\') {
      o += '\\';
      o += (char)c;
    } else if (c < 0x20) {
      char buf[8];
      snprintf(buf, sizeof buf, "\\u%04x", c);
      o += buf;
    } else {
      o += (char)c;
    }
  }
  return o + "\"";
}

// when(at(t_mat), scale((spot(asset) - s0).max(0), one(USD)))
// (src/pricer.rs:158-162, 205-208), as serde JSON.
std::unique_ptr<Con> control_call(const std::string& asset, double strike, double t_mat) {
  char num_t[40], num_k[40];
  snprintf(num_t, sizeof num_t, "%.17g", t_mat);
  snprintf(num_k, sizeof num_k, "%.17g", strike);
  std::string j = std::string("{\"When\":[{\"At\":") + num_t + "},{\"Scale\":[{\"Max\":[{\"Sub\":[{\"Spot\":" +
                  json_string(asset) + "},{\"Const\":" + num_k + "}]},{\"Const\":0.0}]},{\"One\":\"USD\"}]}]}";
  return parse_contract(j.c_str());
}

bool register_overflow(const Error& e) {
  return e.code == KT_ERR_UNSUPPORTED && e.msg.find("registers") != std::string::npos;
}

kt_price_result finish(double price, double sample_std, double std_error, uint64_t n) {
  kt_price_result r{};
  r.price = price;
  r.sample_std = sample_std;
  r.std_error = std_error;
  r.ci95_low = price - kZ95 * std_error;
  r.ci95_high = price + kZ95 * std_error;
  r.n_paths = n;
  return r;
}

}  // namespace

double bs_call_reference(double s, double k, double t, double r, double sigma) { return bs_call_as(s, k, t, r, sigma); }

kt_price_result price_vr(kt_engine* E, const Con& con, const ModelDesc& md, const std::vector<double>& grid,
                         double rate, uint64_t n_paths, uint64_t seed, bool antithetic, bool control_variate) {
  const bool gbm = md.kind == KT_MODEL_GBM;
  // Only the one-asset GBM simulator draws antithetic pairs; the others fall
  // back to plain paths (src/simulator.rs:132-140, 297-314).
  const bool anti = antithetic && gbm;
  if (!control_variate && !anti) {
    if (n_paths == 0) return kt_price_result{};
    Program prog = lower({&con}, grid, md.names);
    return result_from_stats(run_program(E, prog, md, grid, rate, n_paths, seed, nullptr, nullptr)[0]);
  }
  if (!control_variate) {
    // src/variance_reduction.rs:180-203: statistics of the n/2 pair means.
    const uint64_t n_half = n_paths / 2;
    if (n_half == 0) return kt_price_result{};
    Program prog = lower({&con}, grid, md.names);
    auto st = run_program_scen(E, prog, {md}, {rate}, {}, grid, n_half, seed, nullptr, nullptr, nullptr, true);
    return result_from_stats(st[0]);
  }

  // Control: (s0, sigma) of a GBM, (100, 0.2) otherwise (src/pricer.rs:155-163).
  const double s0 = gbm ? md.s0[0] : 100.0;
  const double sigma = gbm ? md.sigma[0] : 0.2;
  const double t_mat = grid.empty() ? 1.0 : grid.back();
  const double bs = bs_call_as(s0, s0, t_mat, rate, sigma);
  auto ctrl = control_call(md.names[0], s0, t_mat);

  if (anti) {
    // src/pricer.rs:190-228: the target's and the control's pair means; the
    // reported spread is the target's, over 2 n_half draws.
    const uint64_t n_half = n_paths / 2;
    KStats t{}, c{};
    if (n_half > 0) {
      try {
        Program prog = lower({&con, ctrl.get()}, grid, md.names);
        auto st = run_program_scen(E, prog, {md}, {rate}, {}, grid, n_half, seed, nullptr, nullptr, nullptr, true);
        t = st[0];
        c = st[1];
      } catch (const Error& e) {
        if (!register_overflow(e)) throw;
        // Same counter-based stream in two launches: the same pairs.
        Program pt = lower({&con}, grid, md.names);
        t = run_program_scen(E, pt, {md}, {rate}, {}, grid, n_half, seed, nullptr, nullptr, nullptr, true)[0];
        Program pc = lower({ctrl.get()}, grid, md.names);
        c = run_program_scen(E, pc, {md}, {rate}, {}, grid, n_half, seed, nullptr, nullptr, nullptr, true)[0];
      }
    }
    const kt_price_result rt = result_from_stats(t), rc = result_from_stats(c);
    const double price = rt.price - 1.0 * (rc.price - bs);
    const uint64_t n = 2 * n_half;
    const double se = rt.sample_std / std::sqrt((double)n);  // 0/0 = NaN when there is no pair
    return finish(price, rt.sample_std, se, n);
  }

  // src/pricer.rs:150-165, src/variance_reduction.rs:210-235: statistics of
  // pv_target - (pv_control - bs) per path.
  if (n_paths == 0) return kt_price_result{};
  KStats d{};
  try {
    Program prog = lower({&con, ctrl.get()}, grid, md.names);
    d = run_program_scen(E, prog, {md}, {rate}, {{1.0, -1.0}}, grid, n_paths, seed, nullptr, nullptr, nullptr, false,
                         2)[0];
  } catch (const Error& e) {
    if (!register_overflow(e)) throw;
    // Two launches keep their per-path values on the device; the difference
    // is reduced there.
    double* a_dev = (double*)E->ensure(E->pv, n_paths * sizeof(double));
    double* b_dev = (double*)E->ensure(E->z, n_paths * sizeof(double));
    Program pt = lower({&con}, grid, md.names);
    run_program(E, pt, md, grid, rate, n_paths, seed, nullptr, nullptr, a_dev);
    Program pc = lower({ctrl.get()}, grid, md.names);
    run_program(E, pc, md, grid, rate, n_paths, seed, nullptr, nullptr, b_dev);
    const int nb = (int)((n_paths + kBlock - 1) / kBlock);
    KStats* parts = (KStats*)E->ensure(E->scratch, (size_t)nb * sizeof(KStats));
    KStats* out = (KStats*)E->ensure(E->stats, 8 * sizeof(KStats));
    check_hip(launch_diff_stats(a_dev, b_dev, 1.0, n_paths, parts, E->stream), "diff stats");
    check_hip(launch_finalize(parts, nb, 1, 1, out, E->stream), "finalize launch");
    check_hip(hipMemcpyAsync(&d, out, sizeof(KStats), hipMemcpyDeviceToHost, E->stream), "copy stats");
    check_hip(hipStreamSynchronize(E->stream), "stream sync");
  }
  kt_price_result r = result_from_stats(d);
  return finish(r.price + bs, r.sample_std, r.std_error, r.n_paths);
}

}  // namespace kt
