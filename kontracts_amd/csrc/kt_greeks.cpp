// kt_greeks.cpp — Greeks by CRN bump-and-reprice (src/greeks.rs:60-98).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <map>
#include <string>

#include "kt_engine_internal.h"

namespace kt {

namespace {

// The seven repricings of greeks_gbm (src/greeks.rs:68-89): base, spot up /
// down, vol up / down, rate up / down (the drift follows the rate).
std::vector<ModelDesc> gbm_scenarios(const char* asset, double s0, double sigma, double r, const kt_bumps& h) {
  auto model = [&](double s, double sg, double rr) {
    ModelDesc md;
    md.kind = KT_MODEL_GBM;
    md.names = {asset};
    md.s0 = {s};
    md.mu = {rr};
    md.sigma = {sg};
    return md;
  };
  return {model(s0, sigma, r),         model(s0 + h.spot, sigma, r), model(s0 - h.spot, sigma, r),
          model(s0, sigma + h.vol, r), model(s0, sigma - h.vol, r),  model(s0, sigma, r + h.rate),
          model(s0, sigma, r - h.rate)};
}

std::vector<double> scenario_rates(double r, const kt_bumps& h) { return {r, r, r, r, r, r + h.rate, r - h.rate}; }

// Output j of the scenario kernel = sum_sc w[j][sc] pv_sc: price, delta,
// gamma, vega, rho as central differences of the seven present values.
std::vector<std::vector<double>> scenario_weights(const kt_bumps& h) {
  const double hs = h.spot, hv = h.vol, hr = h.rate;
  return {{1, 0, 0, 0, 0, 0, 0},
          {0, 1 / (2 * hs), -1 / (2 * hs), 0, 0, 0, 0},
          {-2 / (hs * hs), 1 / (hs * hs), 1 / (hs * hs), 0, 0, 0, 0},
          {0, 0, 0, 1 / (2 * hv), -1 / (2 * hv), 0, 0},
          {0, 0, 0, 0, 0, 1 / (2 * hr), -1 / (2 * hr)}};
}

std::vector<double> greeks_grid(const Con& con, const kt_mc_config* cfg) {
  Plan plan = compile_plan(con);
  return cfg->grid && cfg->n_grid ? std::vector<double>(cfg->grid, cfg->grid + cfg->n_grid)
                                  : time_grid_forced(plan, cfg->steps_per_year, false);
}

void fill_greeks(const std::vector<KStats>& st, size_t first, kt_greeks* out) {
  double* vals[5] = {&out->price, &out->delta, &out->gamma, &out->vega, &out->rho};
  for (int j = 0; j < 5; ++j) {
    kt_price_result pr = result_from_stats(st[first + j]);
    *vals[j] = pr.price;
    out->std_error[j] = pr.std_error;
  }
}

}  // namespace

// greeks_gbm: drift = rate, seven repricings at the same seed.  The seven
// scenarios run in one launch that shares every normal (common random
// numbers); the kernel reduces the per-path samples of price, delta, gamma,
// vega and rho, whose means are exactly the reference's finite differences of
// the seven prices, and whose spread gives each Greek's standard error.
void greeks_gbm_crn(kt_engine* E, const char* json, const char* asset, double s0, double sigma,
                    const kt_mc_config* cfg, const kt_bumps& h, kt_greeks* out) {
  auto con = parse_contract(json);
  std::vector<double> grid = greeks_grid(*con, cfg);
  const double r = cfg->rate;
  std::vector<ModelDesc> scen = gbm_scenarios(asset, s0, sigma, r, h);
  std::vector<double> rates = scenario_rates(r, h);
  for (const ModelDesc& m : scen) validate_model(m);
  validate_grid(grid);
  *out = kt_greeks{};
  const double nan = std::numeric_limits<double>::quiet_NaN();
  if (cfg->antithetic || cfg->control_variate) {
    // Each repricing is a price_gbm with the caller's variance reduction
    // (src/greeks.rs:68-76): seven variance-reduced prices at the same seed.
    double p[7];
    double se0 = 0.0;
    for (int i = 0; i < 7; ++i) {
      kt_price_result r = price_vr(E, *con, scen[i], grid, rates[i], cfg->n_paths, cfg->seed, cfg->antithetic != 0,
                                   cfg->control_variate != 0);
      p[i] = r.price;
      if (i == 0) se0 = r.std_error;
    }
    out->price = p[0];
    out->delta = (p[1] - p[2]) / (2.0 * h.spot);
    out->gamma = (p[1] - 2.0 * p[0] + p[2]) / (h.spot * h.spot);
    out->vega = (p[3] - p[4]) / (2.0 * h.vol);
    out->rho = (p[5] - p[6]) / (2.0 * h.rate);
    out->std_error[0] = se0;
    for (int i = 1; i < 5; ++i) out->std_error[i] = nan;
    return;
  }
  if (cfg->n_paths == 0) return;
  Program prog = lower({con.get()}, grid, {asset});
  int ns = 0, nout = 0;
  if (path_variant_exists(MK_GBM, 1, prog.n_slots, 1, false, kScenarios, &ns, &nout)) {
    auto st = run_program_scen(E, prog, scen, rates, scenario_weights(h), grid, cfg->n_paths, cfg->seed, nullptr,
                               nullptr);
    fill_greeks(st, 0, out);
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
  out->std_error[0] = se0;
  for (int i = 1; i < 5; ++i) out->std_error[i] = nan;
}

// greek_surface (src/surface.rs:104-138): greeks_gbm at every (spot, vol)
// cell, row-major [spot][vol].  Every cell is the seven-scenario CRN launch of
// greeks_gbm at the same seed; all cells run in one grid (blockIdx.y = cell)
// and one finalize, so a cell's result equals greeks_gbm at that point.
// Variance-reduced configurations and programs beyond the scenario kernel go
// cell by cell through greeks_gbm.
void greek_surface(kt_engine* E, const char* json, const char* asset, const double* spots, uint64_t n_spots,
                   const double* vols, uint64_t n_vols, const kt_mc_config* cfg, const kt_bumps& h, kt_greeks* out) {
  if (n_spots == 0 || n_vols == 0) raise(KT_ERR_INCONSISTENT_PATH, "empty surface grid");
  if (!spots || !vols) raise(KT_ERR_INVALID_ARGUMENT, "null surface axis");
  if (n_spots > (1u << 20) || n_vols > (1u << 20) || n_spots * n_vols > (1u << 24))
    raise(KT_ERR_INVALID_ARGUMENT, "surface too large");
  const uint64_t n_cells = n_spots * n_vols;
  auto con = parse_contract(json);
  std::vector<double> grid = greeks_grid(*con, cfg);
  const double r = cfg->rate;
  const std::vector<double> rates = scenario_rates(r, h);
  std::vector<std::vector<ModelDesc>> scen(n_cells);
  for (uint64_t c = 0; c < n_cells; ++c) {
    scen[c] = gbm_scenarios(asset, spots[c / n_vols], vols[c % n_vols], r, h);
    for (const ModelDesc& m : scen[c]) validate_model(m);
  }
  validate_grid(grid);
  for (uint64_t c = 0; c < n_cells; ++c) out[c] = kt_greeks{};
  Program prog = lower({con.get()}, grid, {asset});
  const uint64_t N = cfg->n_paths;
  const uint64_t nb = (N + kBlock - 1) / kBlock;
  int ns = 0, nout = 0;
  const bool fast = !cfg->antithetic && !cfg->control_variate && N > 0 && nb <= (1u << 20) &&
                    path_variant_exists(MK_GBM, 1, prog.n_slots, 1, false, kScenarios, &ns, &nout);
  if (!fast) {
    for (uint64_t c = 0; c < n_cells; ++c)
      greeks_gbm_crn(E, json, asset, spots[c / n_vols], vols[c % n_vols], cfg, h, &out[c]);
    return;
  }
  std::vector<KOp> ops = prog.ops;
  if (ops.empty()) ops.push_back(KOp{OP_NOP, 0, 0, 0, 0.0f});
  const std::vector<std::vector<double>> w = scenario_weights(h);
  auto a256 = [](size_t b) { return (b + 255) & ~(size_t)255; };
  // Cells run vol-major: the model table depends on the vol only and the
  // thresholds on the spot only, so identical tables are stored once and the
  // cells resident together mostly read one table.  slot p of a launch is
  // cell order[p].
  std::vector<uint64_t> order;
  order.reserve(n_cells);
  for (uint64_t j = 0; j < n_vols; ++j)
    for (uint64_t i = 0; i < n_spots; ++i) order.push_back(i * n_vols + j);
  // Cells per launch: ~2 GB of partials at most.
  const uint64_t per_launch = std::max<uint64_t>(1, (1ull << 24) / nb);
  for (uint64_t c0 = 0; c0 < n_cells; c0 += per_launch) {
    const uint64_t nc = std::min(per_launch, n_cells - c0);
    bool serial = true;  // every cell takes the scenario-serial kernel
    std::vector<LaunchTables> T(nc);
    // the discount tables depend on the rates only: the first cell's serve all
    for (uint64_t i = 0; i < nc; ++i) T[i] = launch_tables(prog, scen[order[c0 + i]], rates, grid, i == 0);
    // Arena: [cells' PathArgs][ops][steps][init][df][distinct groups / tab / thr blobs]
    size_t off = a256(nc * sizeof(PathArgs));
    const size_t o_ops = off;
    off += a256(ops.size() * sizeof(KOp));
    const size_t o_steps = off;
    off += a256(prog.steps.size() * sizeof(KStep));
    const size_t o_init = off;
    off += a256(std::max<size_t>(prog.init.size(), 1) * sizeof(KOp));
    const size_t o_df = off;
    off += a256(T[0].df.size() * sizeof(double));
    std::map<std::string, size_t> interned;
    std::vector<std::pair<const void*, size_t>> blobs;  // (data, bytes) at offsets blob_off
    std::vector<size_t> blob_off;
    auto intern = [&](const void* p, size_t bytes) {
      std::string key((const char*)p, bytes);
      auto it = interned.find(key);
      if (it != interned.end()) return it->second;
      const size_t o = off;
      off += a256(std::max<size_t>(bytes, 1));
      interned.emplace(std::move(key), o);
      blobs.push_back({p, bytes});
      blob_off.push_back(o);
      return o;
    };
    std::vector<size_t> o_grp(nc), o_tab(nc), o_thr(nc);
    for (uint64_t i = 0; i < nc; ++i) {
      o_grp[i] = intern(T[i].groups.data(), T[i].groups.size() * sizeof(KGroup));
      o_tab[i] = intern(T[i].tab.data(), T[i].tab.size() * sizeof(float));
      o_thr[i] = intern(T[i].thr.data(), T[i].thr.size() * sizeof(float));
    }
    char* dev = (char*)E->ensure(E->arena, off);
    KStats* partials = (KStats*)E->ensure(E->partials, nc * nb * kScenarioStats * sizeof(KStats));
    KStats* stats = (KStats*)E->ensure(E->stats, nc * kScenarioStats * sizeof(KStats));
    unsigned* err = (unsigned*)E->ensure(E->errw, 256);
    char* st = E->stage(off);
    memset(st, 0, off);
    auto put = [&](size_t o, const void* p, size_t bytes) {
      if (bytes) memcpy(st + o, p, bytes);
    };
    put(o_ops, ops.data(), ops.size() * sizeof(KOp));
    put(o_steps, prog.steps.data(), prog.steps.size() * sizeof(KStep));
    put(o_init, prog.init.data(), prog.init.size() * sizeof(KOp));
    put(o_df, T[0].df.data(), T[0].df.size() * sizeof(double));
    for (size_t b = 0; b < blobs.size(); ++b) put(blob_off[b], blobs[b].first, blobs[b].second);
    for (uint64_t i = 0; i < nc; ++i) {
      const std::vector<ModelDesc>& sc_models = scen[order[c0 + i]];
      PathArgs A{};
      A.ops = (const KOp*)(dev + o_ops);
      A.steps = (const KStep*)(dev + o_steps);
      A.groups = (const KGroup*)(dev + o_grp[i]);
      A.init = (const KOp*)(dev + o_init);
      A.df = (const double*)(dev + o_df);
      A.tab = (const float*)(dev + o_tab[i]);
      A.thr = (const float*)(dev + o_thr[i]);
      A.n_init = (int)prog.init.size();
      A.n_grid = (int)grid.size();
      A.n_groups = (int)T[i].groups.size();
      A.n_uniform = uniform_group_prefix(T[i].groups.data(), A.n_groups);
      A.serial_tail = scenario_serial_tail(T[i].groups.data(), A.n_groups);
      serial = serial && A.serial_tail;
      A.tab_uniform = uniform_table_rows(T[i].tab, grid.size(), (size_t)T[i].stride);
      A.tab_stride = T[i].stride;
      fill_model_args(A, sc_models[0], cfg->seed);
      for (int sc = 0; sc < kScenarios; ++sc) {
        A.s0_scen[sc] = (float)sc_models[sc].s0[0];
        A.df_index[sc] = T[i].df_index[sc];
      }
      for (int j = 0; j < kScenarioStats; ++j)
        for (int sc = 0; sc < kScenarios; ++sc) A.stat_w[j][sc] = w[j][sc];
      A.n_paths = N;
      A.path_base = 0;
      A.partials = partials + i * nb * kScenarioStats;
      A.err = err;
      memcpy(st + i * sizeof(PathArgs), &A, sizeof(PathArgs));
    }
    E->stage_copy(dev, off);
    check_hip(hipMemsetAsync(err, 0, sizeof(unsigned), E->stream), "memset");
    E->prof_begin();
    int n_cu = 0;
    check_hip(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, E->device), "device attribute");
    int& occ = E->occupancy_surface[(ns == 8 ? 0 : 1) + (serial ? 2 : 0)];
    if (occ <= 0) occ = surface_occupancy_blocks_per_cu(ns, serial);
    check_hip(launch_surface(ns, serial, (const PathArgs*)dev, (int)nc, (int)nb, partials, stats, n_cu, occ,
                             E->stream),
              "surface launch");
    E->prof_end((double)N * (double)(grid.size() - 1) * kScenarios * (double)nc);
    std::vector<KStats> res(nc * kScenarioStats);
    unsigned err_bits = 0;
    check_hip(hipMemcpyAsync(res.data(), stats, res.size() * sizeof(KStats), hipMemcpyDeviceToHost, E->stream),
              "copy stats");
    check_hip(hipMemcpyAsync(&err_bits, err, sizeof(unsigned), hipMemcpyDeviceToHost, E->stream), "copy err");
    check_hip(hipStreamSynchronize(E->stream), "stream sync");
    if (err_bits) {
      int code = lowest_error_code(err_bits);
      raise(code, error_text(code, prog));
    }
    for (uint64_t i = 0; i < nc; ++i) fill_greeks(res, i * kScenarioStats, &out[order[c0 + i]]);
  }
}

// Central differences over the model parameters with common random numbers.
// With stderr_out, each bumped pair keeps its per-path present values on the
// device and the standard error of the per-path differences is reduced there.
void greeks_bump(kt_engine* E, const char* json, const kt_model* model, const kt_mc_config* cfg, const double* h,
                 uint64_t n_params, double* out, double* stderr_out) {
  ModelDesc base = model_from_abi(model);
  PreparedContract pc = prepare(json, base, cfg);
  uint64_t want = base.kind == KT_MODEL_HESTON ? 6 : (base.kind == KT_MODEL_GBM ? 3 : 0);
  if (want == 0) raise(KT_ERR_UNSUPPORTED, "bump Greeks are defined for GBM and Heston models");
  if (n_params != want) raise(KT_ERR_INVALID_ARGUMENT, "wrong number of bump sizes for this model");
  if (!h) raise(KT_ERR_INVALID_ARGUMENT, "null bump sizes");
  for (uint64_t i = 0; i <= n_params; ++i) {
    out[i] = 0.0;
    if (stderr_out) stderr_out[i] = 0.0;
  }
  const uint64_t N = cfg->n_paths;
  if (N == 0) return;
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
    validate_model(m);
    return m;
  };
  Program prog = lower({pc.con.get()}, pc.grid, base.names);
  auto st0 = run_program(E, prog, base, pc.grid, cfg->rate, N, cfg->seed, nullptr, nullptr);
  kt_price_result r0 = result_from_stats(st0[0]);
  out[0] = r0.price;
  if (stderr_out) stderr_out[0] = r0.std_error;
  double* up_dev = nullptr;
  double* dn_dev = nullptr;
  KStats* diff_partials = nullptr;
  KStats* diff_stats = nullptr;
  if (stderr_out) {
    up_dev = (double*)E->ensure(E->pv, N * sizeof(double));
    dn_dev = (double*)E->ensure(E->z, N * sizeof(double));
    diff_partials = (KStats*)E->ensure(E->scratch, ((N + 255) / 256) * sizeof(KStats));
    diff_stats = (KStats*)E->ensure(E->stats, kAdjMaxOut * sizeof(KStats));
  }
  for (uint64_t j = 0; j < n_params; ++j) {
    auto su = run_program(E, prog, bumped((int)j, h[j]), pc.grid, cfg->rate, N, cfg->seed, nullptr, nullptr, up_dev);
    auto sd = run_program(E, prog, bumped((int)j, -h[j]), pc.grid, cfg->rate, N, cfg->seed, nullptr, nullptr, dn_dev);
    out[1 + j] = (result_from_stats(su[0]).price - result_from_stats(sd[0]).price) / (2.0 * h[j]);
    if (stderr_out) {
      const int nb = (int)((N + 255) / 256);
      check_hip(launch_diff_stats(up_dev, dn_dev, 1.0 / (2.0 * h[j]), N, diff_partials, E->stream), "diff stats");
      check_hip(launch_finalize(diff_partials, nb, 1, 1, diff_stats, E->stream), "finalize launch");
      KStats ds{};
      check_hip(hipMemcpyAsync(&ds, diff_stats, sizeof(KStats), hipMemcpyDeviceToHost, E->stream), "copy stats");
      check_hip(hipStreamSynchronize(E->stream), "stream sync");
      stderr_out[1 + j] = result_from_stats(ds).std_error;
    }
  }
}

// Pathwise Greeks (kt_adjoint.hip): one launch; per-path samples of the
// price and of each parameter derivative, reduced in FP64 like prices.
AdjJobHost adjoint_job_host(const Program& prog, const ModelDesc& md, const std::vector<double>& grid, double rate,
                            AdjPlan plan) {
  AdjJobHost H;
  H.tab = model_table(md, grid, &H.stride);
  H.df.resize(grid.size());
  discount_factors(prog, grid, rate, H.df.data());
  // GBM / basket: {sqrt(dt), t log2(e)} per point (d x / d sigma is
  // (W_t - sigma t) log2(e), W_t the sum of sqrt(dt) w; d x / d mu is t log2(e))
  H.adj.assign(2 * grid.size(), 0.0f);
  std::vector<double> dts(grid.size(), 0.0);
  double t_acc = 0.0;
  for (size_t k = 1; k < grid.size(); ++k) {
    const double dt = grid[k] - grid[k - 1];
    if (dt > 0.0) {
      dts[k] = dt;
      t_acc += dt;
      H.adj[2 * k] = (float)std::sqrt(dt);
    }
    H.adj[2 * k + 1] = (float)(t_acc * kLog2e);
  }
  H.ops = prog.ops;
  if (H.ops.empty()) H.ops.push_back(KOp{OP_NOP, 0, 0, 0, 0.0f});
  H.thr.assign(1, 0.0f);
  H.plan = std::move(plan);
  H.rev.assign(8, 0.0f);
  if (md.kind == KT_MODEL_HESTON) {
    H.rev.assign(8 * grid.size(), 0.0f);
    for (size_t k = 0; k < grid.size(); ++k) {
      const float* t = &H.tab[k * kHestonStride];
      float* r = &H.rev[8 * k];
      r[0] = t[1];
      r[1] = t[2];
      r[2] = t[3];
      r[3] = t[5];
      r[4] = (float)dts[k];
      r[5] = (float)std::sqrt(dts[k]);
      r[6] = (float)md.kappa * (float)dts[k];
    }
  }
  return H;
}

std::vector<std::pair<const void*, size_t>> adjoint_blobs(const AdjJobHost& H, const Program& prog) {
  return {{H.ops.data(), H.ops.size() * sizeof(KOp)},
          {prog.steps.data(), prog.steps.size() * sizeof(KStep)},
          {prog.groups.data(), prog.groups.size() * sizeof(KGroup)},
          {prog.init.data(), prog.init.size() * sizeof(KOp)},
          {H.df.data(), H.df.size() * sizeof(double)},
          {H.tab.data(), H.tab.size() * sizeof(float)},
          {H.thr.data(), H.thr.size() * sizeof(float)},
          {H.adj.data(), H.adj.size() * sizeof(float)},
          {H.plan.feed.data(), H.plan.feed.size() * sizeof(uint32_t)},
          {H.rev.data(), H.rev.size() * sizeof(float)}};
}

void fill_adjoint_args(AdjArgs& A, const AdjJobHost& H, const Program& prog, const ModelDesc& md,
                       const std::vector<double>& grid, uint64_t seed, const std::vector<void*>& dev) {
  PathArgs& P = A.p;
  P.ops = (const KOp*)dev[0];
  P.steps = (const KStep*)dev[1];
  P.groups = (const KGroup*)dev[2];
  P.init = (const KOp*)dev[3];
  P.df = (const double*)dev[4];
  P.tab = (const float*)dev[5];
  P.thr = (const float*)dev[6];
  P.n_init = (int)prog.init.size();
  P.n_grid = (int)grid.size();
  P.n_groups = (int)((grid.size() + 3) / 4);
  // leading groups whose points run no ops: the Heston kernels advance them
  // with no interpretation
  P.n_uniform = !prog.groups.empty() && prog.groups[0].kind == G_NONE
                    ? std::min(uniform_group_prefix(prog.groups.data(), (int)prog.groups.size()), P.n_groups)
                    : 0;
  P.tab_stride = H.stride;
  // the Heston sweep, and the GBM / basket runs of op-free groups, keep the
  // shared rows in registers when every row is equal (GBM and basket: the
  // model rows and sqrt(dt); Heston: the model and reverse rows)
  bool sq_uniform = grid.size() >= 3;
  for (size_t k = 2; k < grid.size() && sq_uniform; ++k) sq_uniform = H.adj[2 * k] == H.adj[2];
  P.tab_uniform = uniform_table_rows(H.tab, grid.size(), (size_t)H.stride) &&
                  (md.kind == KT_MODEL_HESTON ? uniform_table_rows(H.rev, grid.size(), 8) != 0 : sq_uniform);
  fill_model_args(P, md, seed);
  A.adj_tab = (const float*)dev[7];
  A.feed = (const uint32_t*)dev[8];
  A.rev = (const float*)dev[9];
  A.n_acc = H.plan.n_acc;
  for (int q = 0; q < 4; ++q) {
    A.acc_slot[q] = H.plan.slot[q];
    A.acc_kind[q] = H.plan.kind[q];
    A.acc_asset[q] = H.plan.asset[q];
  }
  const int K = (int)md.names.size();
  for (int a = 0; a < K && a < kMaxAssets; ++a) {
    A.sigma[a] = (float)md.sigma[a];
    A.mu[a] = (float)md.mu[a];
  }
  A.kappa = (float)md.kappa;
  A.theta = (float)md.theta;
  A.sigma_v = (float)md.sigma_v;
  A.seg = kAdjSeg;
}

void greeks_adjoint(kt_engine* E, const char* json, const kt_model* model, const kt_mc_config* cfg, double* out,
                    double* stderr_out, uint64_t n_out, std::vector<KStats>* stats_out) {
  ModelDesc md = model_from_abi(model);
  PreparedContract pc = prepare(json, md, cfg);
  const int kind = model_kernel_kind(md);
  const int K = (int)md.names.size();
  const int need = adjoint_outputs(kind, K);
  if (n_out < (uint64_t)need)
    raise(KT_ERR_INVALID_ARGUMENT, "output arrays need " + std::to_string(need) + " entries for this model");
  const std::vector<double>& grid = pc.grid;
  Program prog = lower({pc.con.get()}, grid, md.names);
  bool flags = false;
  for (const KOp& op : prog.ops) flags = flags || op.code == OP_FLAGERR;
  if (flags) {
    // surface the path error the pricer raises (unknown asset, `or` reached, ...)
    run_program(E, prog, md, grid, cfg->rate, cfg->n_paths, cfg->seed, nullptr, nullptr);
  }
  AdjPlan plan = classify_adjoint(prog);
  int ns = 0;
  if (!adjoint_variant_exists(kind, K, prog.n_slots, &ns))
    raise(KT_ERR_UNSUPPORTED, "pathwise Greeks: program needs more registers than the kernel holds");
  for (int j = 0; j < need; ++j) {
    out[j] = 0.0;
    if (stderr_out) stderr_out[j] = 0.0;
  }
  const uint64_t N = cfg->n_paths;
  if (N == 0) return;

  AdjJobHost H = adjoint_job_host(prog, md, grid, cfg->rate, std::move(plan));
  auto dev = E->upload(adjoint_blobs(H, prog));

  const uint64_t total_blocks = (N + kBlock - 1) / kBlock;
  if (total_blocks > (1ull << 31)) raise(KT_ERR_INVALID_ARGUMENT, "too many paths");
  KStats* partials = (KStats*)E->ensure(E->partials, total_blocks * need * sizeof(KStats));
  KStats* stats = (KStats*)E->ensure(E->stats, kAdjMaxOut * sizeof(KStats));
  float* ck = nullptr;
  const int seg = E->adj_seg;
  const size_t n_ck = (grid.size() - 1) / seg + 1;
  const bool tangent = kind == MK_HESTON && E->pathwise_tangent;
  if (kind == MK_HESTON && !tangent) ck = (float*)E->ensure(E->aux, n_ck * 2 * N * sizeof(float));

  AdjArgs A{};
  fill_adjoint_args(A, H, prog, md, grid, cfg->seed, dev);
  PathArgs& P = A.p;
  A.ck = ck;
  A.ck_stride = N;
  A.seg = seg;
  A.tangent = tangent ? 1 : 0;

  E->prof_begin();
  const uint64_t kMaxBlocksPerLaunch = 1ull << 20;
  for (uint64_t b0 = 0; b0 < total_blocks; b0 += kMaxBlocksPerLaunch) {
    const uint64_t nb = std::min(kMaxBlocksPerLaunch, total_blocks - b0);
    P.path_base = b0 * kBlock;
    P.n_paths = std::min<uint64_t>(nb * kBlock, N - b0 * kBlock);
    P.partials = partials + b0 * need;
    check_hip(launch_adjoint(kind, K, ns, A, (int)nb, E->stream), "adjoint kernel launch");
  }
  E->prof_end((double)N * (double)(grid.size() - 1));
  check_hip(launch_finalize(partials, (int)total_blocks, need, need, stats, E->stream), "finalize launch");
  std::vector<KStats> st(need);
  check_hip(hipMemcpyAsync(st.data(), stats, need * sizeof(KStats), hipMemcpyDeviceToHost, E->stream), "copy stats");
  check_hip(hipStreamSynchronize(E->stream), "stream sync");
  for (int j = 0; j < need; ++j) {
    kt_price_result r = result_from_stats(st[j]);
    out[j] = r.price;
    if (stderr_out) stderr_out[j] = r.std_error;
  }
  if (stats_out) *stats_out = st;
}

}  // namespace kt
