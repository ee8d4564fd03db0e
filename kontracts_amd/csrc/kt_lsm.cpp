// kt_lsm.cpp — American pricing by Longstaff-Schwartz (src/lsm.rs:80-229).
//
// Device pipeline for D exercise dates and N paths:
//   1. lsm_paths_kernel: simulate [0, dates...] once and store the exercised
//      payoff and the regression spot per (date, path) as an FP32 pair,
//      [date][path].
//   2. D passes from the last date back: pass d applies the exercise decision
//      of date d and accumulates date d-1's ITM normal equations in FP64 per
//      block; a one-block kernel reduces the block partials in a fixed order
//      and solves the m x m system for the next pass.  The last pass (d = 0)
//      discounts to t = 0 and reduces the price statistics.
// Everything stays on the engine's stream; the host waits once at the end.
#include <algorithm>
#include <cmath>

#include "kt_engine_internal.h"

namespace kt {

namespace {

// The exercised payoff as a vanilla put / call, when every exercise date's
// op list is spot(asset 0) -> subi / rsubi strike -> maxi floor -> emit of
// that times the constant-1 register with a scale: the paths kernel then
// evaluates it as straight-line code (the same operations).
LsmPayoff lsm_vanilla_payoff(const Program& prog, int model_kind) {
  LsmPayoff none{0, 0.0f, 0.0f, 0.0f};
  if (model_kind != MK_GBM && model_kind != MK_HESTON) return none;
  bool one = false;
  for (const KOp& op : prog.init) {
    if (op.d != kSlotOne) continue;
    one = op.code == OP_CONST && op.imm == 1.0f;
  }
  if (!one || prog.steps.size() < 2 || prog.steps[0].count != 0) return none;
  LsmPayoff pay = none;
  for (size_t k = 1; k < prog.steps.size(); ++k) {
    const KStep& st = prog.steps[k];
    if (st.count != 4 || (size_t)st.begin + 4 > prog.ops.size()) return none;
    const KOp* o = &prog.ops[st.begin];
    if (o[0].code != OP_SPOT || o[0].a != 0) return none;
    if ((o[1].code != OP_SUBI && o[1].code != OP_RSUBI) || o[1].a != o[0].d) return none;
    if (o[2].code != OP_MAXI || o[2].a != o[1].d) return none;
    if (o[3].code != OP_EMIT || o[3].a != o[2].d || o[3].b != kSlotOne) return none;
    if (o[0].d == kSlotOne || o[1].d == kSlotOne || o[2].d == kSlotOne) return none;
    const LsmPayoff here{o[1].code == OP_SUBI ? 1 : 2, o[1].imm, o[2].imm, o[3].imm};
    if (k > 1 && (here.kind != pay.kind || here.strike != pay.strike || here.floor != pay.floor ||
                  here.scale != pay.scale))
      return none;
    pay = here;
  }
  return pay;
}

}  // namespace

void price_american(kt_engine* E, const char* json, const double* dates, uint64_t n_dates, const kt_model* model,
                    const kt_mc_config* cfg, int n_basis, kt_price_result* out, double* pv_host,
                    double* coef_host) {
  if (!cfg || !out) raise(KT_ERR_INVALID_ARGUMENT, "null config or output");
  auto con = parse_contract(json);
  // Argument checks in the order of src/lsm.rs:87-101.
  if (n_dates == 0 || !dates) raise(KT_ERR_MALFORMED_CONTRACT, "LSM: at least one exercise date is required");
  for (uint64_t i = 1; i < n_dates; ++i)
    if (!(dates[i] > dates[i - 1]))
      raise(KT_ERR_MALFORMED_CONTRACT, "LSM: exercise dates must be strictly increasing and > 0");
  if (!(dates[0] > 0.0)) raise(KT_ERR_MALFORMED_CONTRACT, "LSM: exercise dates must be strictly increasing and > 0");
  if (n_basis < 2) raise(KT_ERR_MALFORMED_CONTRACT, "LSM: n_basis >= 2 is required (constant + linear)");
  if (n_basis > kLsmBasisMax)
    raise(KT_ERR_UNSUPPORTED, "LSM: the GPU regression holds at most " + std::to_string(kLsmBasisMax) +
                                  " basis functions");
  ModelDesc md = model_from_abi(model);
  validate_model(md);
  std::vector<double> grid{0.0};
  grid.insert(grid.end(), dates, dates + n_dates);
  validate_grid(grid);
  Program prog = lower_exercise(*con, grid, md.names);

  *out = kt_price_result{};
  const uint64_t N = cfg->n_paths;
  if (N == 0) {
    *out = result_from_stats(KStats{0.0, 0.0, 0.0});
    return;
  }
  const int D = (int)n_dates;
  const int m = n_basis;
  const double rate = cfg->rate;
  const int kind = model_kernel_kind(md);
  const int K = (int)md.names.size();
  const int ns = prog.n_slots <= 16 && kind != MK_BASKET ? 16 : 32;
  if (prog.n_slots > ns) raise(KT_ERR_UNSUPPORTED, "exercised payoff needs more registers than the kernel holds");

  int stride = 0;
  std::vector<float> tab = model_table(md, grid, &stride);
  std::vector<double> df(grid.size(), 1.0);  // flows are valued at the exercise date
  std::vector<KOp> ops = prog.ops;
  if (ops.empty()) ops.push_back(KOp{OP_NOP, 0, 0, 0, 0.0f});
  std::vector<float> thr(1, 0.0f);
  auto dev = E->upload({{ops.data(), ops.size() * sizeof(KOp)},
                        {prog.steps.data(), prog.steps.size() * sizeof(KStep)},
                        {prog.groups.data(), prog.groups.size() * sizeof(KGroup)},
                        {prog.init.data(), prog.init.size() * sizeof(KOp)},
                        {df.data(), df.size() * sizeof(double)},
                        {tab.data(), tab.size() * sizeof(float)},
                        {thr.data(), thr.size() * sizeof(float)}});

  const uint64_t total_blocks = (N + kBlock - 1) / kBlock;
  if (total_blocks > (1ull << 31) - 1) raise(KT_ERR_INVALID_ARGUMENT, "too many paths");
  const int nbp = (int)std::min<uint64_t>(total_blocks, kLsmPassBlocks);
  const size_t n_cols = (size_t)D * N;
  float2* cols = (float2*)E->ensure(E->paths, n_cols * sizeof(float2));
  auto align = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t value_bytes = align(N * sizeof(double));
  const size_t reg_bytes = align((size_t)kLsmRegCount * nbp * sizeof(double));
  const size_t coef_rows = (size_t)std::max(D - 1, 1);
  const size_t coef_bytes = align(coef_rows * (kLsmBasisMax + 1) * sizeof(double));
  char* aux = (char*)E->ensure(E->aux, value_bytes + reg_bytes + coef_bytes);
  double* value = (double*)aux;
  double* reg = (double*)(aux + value_bytes);
  double* coef = (double*)(aux + value_bytes + reg_bytes);
  KStats* partials = (KStats*)E->ensure(E->partials, (size_t)nbp * sizeof(KStats));
  KStats* stats = (KStats*)E->ensure(E->stats, 8 * sizeof(KStats));
  unsigned* err = (unsigned*)E->ensure(E->errw, 256);
  check_hip(hipMemsetAsync(err, 0, sizeof(unsigned), E->stream), "memset");

  PathArgs A{};
  A.ops = (const KOp*)dev[0];
  A.steps = (const KStep*)dev[1];
  A.groups = (const KGroup*)dev[2];
  A.init = (const KOp*)dev[3];
  A.df = (const double*)dev[4];
  A.tab = (const float*)dev[5];
  A.thr = (const float*)dev[6];
  A.n_init = (int)prog.init.size();
  A.n_grid = (int)grid.size();
  A.n_groups = (int)prog.groups.size();
  A.n_uniform = uniform_group_prefix(prog.groups.data(), A.n_groups);
  A.tab_stride = stride;
  fill_model_args(A, md, cfg->seed);
  A.err = err;

  const LsmPayoff pay = lsm_vanilla_payoff(prog, kind);
  const uint64_t kMaxBlocksPerLaunch = 1ull << 20;
  auto disc = [&](int d_from, int d_to) {  // exercise-date indices, d_from < d_to
    return std::exp(-rate * (grid[d_to + 1] - grid[d_from + 1]));
  };
  auto coef_row = [&](int d) { return coef + (size_t)d * (kLsmBasisMax + 1); };
  // Pass for date d: decide at d (d = D-1: start from the last payoff), then
  // accumulate date d-1's regression, solved before the next pass; the last
  // pass (d = 0) discounts to t = 0.
  std::vector<LsmPassArgs> passes;
  for (int d = D - 1; d >= 0; --d) {
    LsmPassArgs P{};
    P.cols = cols;
    P.value = value;
    P.reg = reg;
    P.stats = partials;
    P.n = N;
    P.n_dates = D;
    P.m = m;
    P.disc0 = std::exp(-rate * grid[1]);
    P.d_b = d == D - 1 ? -1 : d;
    P.coef = coef_row(d == D - 1 ? 0 : d);  // unused on the first pass
    P.disc_b = d == D - 1 ? 0.0 : disc(d, d + 1);
    P.d_a = d - 1;
    P.disc_a = d >= 1 ? disc(d - 1, d) : 0.0;
    P.final_pass = d == 0;
    passes.push_back(P);
  }
  // The whole dependent chain, from the paths kernel to the finalize.
  auto enqueue = [&]() -> hipError_t {
    for (uint64_t b0 = 0; b0 < total_blocks; b0 += kMaxBlocksPerLaunch) {
      const uint64_t nb = std::min(kMaxBlocksPerLaunch, total_blocks - b0);
      PathArgs Ab = A;
      Ab.path_base = b0 * kBlock;
      Ab.n_paths = std::min<uint64_t>(nb * kBlock, N - b0 * kBlock);
      hipError_t e = launch_lsm_paths(kind, K, ns, Ab, pay, (int)nb, cols + b0 * kBlock, N, E->stream);
      if (e != hipSuccess) return e;
    }
    for (int i = 0; i < D; ++i) {
      hipError_t e = launch_lsm_pass(passes[i], nbp, E->stream);
      if (e != hipSuccess) return e;
      const int d = D - 1 - i;
      if (d >= 1) {
        e = launch_lsm_solve(reg, nbp, m, coef_row(d - 1), E->stream);
        if (e != hipSuccess) return e;
      }
    }
    return launch_finalize(partials, nbp, 1, 1, stats, E->stream);
  };
  E->prof_begin();
  if (!E->lsm_use_graph) {
    check_hip(enqueue(), "LSM launch");
  } else {
    // the graph replays while every launch argument is the same
    std::string key((const char*)&A, sizeof(A));
    key.append((const char*)&pay, sizeof(pay));
    for (const LsmPassArgs& P : passes) key.append((const char*)&P, sizeof(P));
    const uint64_t shape[6] = {N, (uint64_t)D, (uint64_t)m, (uint64_t)nbp, (uint64_t)(uintptr_t)stats,
                               (uint64_t)kind * 64 + (uint64_t)K * 4 + (uint64_t)(ns == 16)};
    key.append((const char*)shape, sizeof(shape));
    if (!E->lsm_exec || E->lsm_key != key) {
      if (E->lsm_exec) (void)hipGraphExecDestroy(E->lsm_exec);
      if (E->lsm_graph) (void)hipGraphDestroy(E->lsm_graph);
      E->lsm_exec = nullptr;
      E->lsm_graph = nullptr;
      E->lsm_key.clear();
      check_hip(hipStreamBeginCapture(E->stream, hipStreamCaptureModeThreadLocal), "begin capture");
      const hipError_t le = enqueue();
      hipGraph_t g = nullptr;
      const hipError_t ce = hipStreamEndCapture(E->stream, &g);
      if (g && (le != hipSuccess || ce != hipSuccess)) (void)hipGraphDestroy(g);
      check_hip(le, "LSM launch");
      check_hip(ce, "end capture");
      E->lsm_graph = g;
      check_hip(hipGraphInstantiate(&E->lsm_exec, g, nullptr, nullptr, 0), "graph instantiate");
      E->lsm_key = key;
    }
    check_hip(hipGraphLaunch(E->lsm_exec, E->stream), "graph launch");
  }
  E->prof_end((double)N * (double)D);

  KStats st{};
  unsigned err_bits = 0;
  check_hip(hipMemcpyAsync(&st, stats, sizeof(KStats), hipMemcpyDeviceToHost, E->stream), "copy stats");
  check_hip(hipMemcpyAsync(&err_bits, err, sizeof(unsigned), hipMemcpyDeviceToHost, E->stream), "copy err");
  // The last pass leaves each path's value discounted to t = 0 in `value`.
  if (pv_host)
    check_hip(hipMemcpyAsync(pv_host, value, N * sizeof(double), hipMemcpyDeviceToHost, E->stream), "copy pv");
  std::vector<double> c_host;
  if (coef_host && D >= 2) {
    c_host.resize((size_t)(D - 1) * (kLsmBasisMax + 1));
    check_hip(hipMemcpyAsync(c_host.data(), coef, c_host.size() * sizeof(double), hipMemcpyDeviceToHost, E->stream),
              "copy coefficients");
  }
  check_hip(hipStreamSynchronize(E->stream), "stream sync");
  if (err_bits) {
    int code = lowest_error_code(err_bits);
    raise(code, error_text(code, prog));
  }
  for (int d = 0; d + 1 < D && coef_host; ++d) {
    for (int j = 0; j < m; ++j) coef_host[(size_t)d * (m + 1) + j] = c_host[(size_t)d * (kLsmBasisMax + 1) + j];
    coef_host[(size_t)d * (m + 1) + m] = c_host[(size_t)d * (kLsmBasisMax + 1) + kLsmBasisMax];
  }
  *out = result_from_stats(st);
}

}  // namespace kt
