// kt_launch.h — kernel arguments and launch entry points shared by the engine
// (host) and the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kt_program.h"

namespace kt {

enum ModelKind { MK_GBM = 0, MK_HESTON = 1, MK_BASKET = 2 };

// Per grid point model tables (floats, tab_stride per point):
//   GBM, basket: for each asset (a, b): log2(S) += a + b * w, where
//     a = (mu - sigma^2/2) dt log2(e), b = sigma sqrt(dt) log2(e)   (0 if dt <= 0)
//   Heston: {c0, c1, c2, e0, d1, d2, 0, 0}:
//     x += c0 + c1 vp + c2 sqrt(vp) z1          (x = log2(S/S0))
//     v' = v + e0 vp + d1 + d2 sqrt(vp) (rho z1 + rho' z2)
//     c0 = r dt log2e, c1 = -dt log2e / 2, c2 = sqrt(dt) log2e,
//     e0 = -kappa dt, d1 = kappa theta dt, d2 = sigma_v sqrt(dt)
constexpr int kHestonStride = 8;

// CRN scenario kernel (greeks_gbm): base, S+h, S-h, sigma+h, sigma-h, r+h, r-h,
// reduced into price, delta, gamma, vega and rho samples.
constexpr int kScenarios = 7;
constexpr int kScenarioStats = 5;

struct PathArgs {
  const KOp* ops;
  const KStep* steps;
  const KGroup* groups;  // [n_groups]
  const KOp* init;
  const double* df;   // [n_df][n_grid] discount factors exp(-r t_k); scenario sc uses table df_index[sc]
  const float* tab;   // [n_grid][tab_stride]
  const float* thr;   // [n_groups][SC] knock-out thresholds log2(H / s0) (knock-out groups only)
  int n_init;
  int n_grid;
  int n_groups;
  int n_uniform;  // leading groups of one fast kind (uniform_group_prefix)
  int tab_uniform;  // rows 1 .. n_grid - 1 of tab are identical (a grid of equal steps)
  int tab_stride;
  int heston_full_truncation;
  uint64_t n_paths;    // paths in this launch
  uint64_t path_base;  // global index of the launch's first path
  uint32_t seed_lo, seed_hi;
  float s0[kMaxAssets];
  float chol[kMaxAssets * kMaxAssets];  // lower-triangular, row-major K x K
  float v0, rho, rho_perp;
  float s0_scen[kScenarios];      // scenario kernel: spot per scenario
  int df_index[kScenarios];       // discount table of each scenario (0 when SC == 1)
  double stat_w[kScenarioStats][kScenarios];  // scenario kernel: output j = sum_sc w[j][sc] pv_sc
  int combine;  // one scenario: > 0 makes output 0 sum_{o < combine} stat_w[0][o] pv_o
  int serial_tail;  // scenarios: scenario_serial_tail of the groups (picks the kernel variant)
  KStats* partials;  // [n_blocks][NOUT]
  unsigned* err;     // bit e set <=> error code e raised on a live path
  double* pv_out;    // debug: per-path pv of output 0 (nullable)
  float* z_out;      // debug: normals [path][grid point][component] (nullable)
};

constexpr int kBlock = 256;

// Path kernel variants available: returns false if no variant matches.
// anti: each lane prices an antithetic pair (GBM, one scenario, no debug
// output); n_paths then counts pairs.
bool path_variant_exists(int model, int k_assets, int n_slots, int n_outputs, bool debug, int n_scen,
                         int* ns_chosen, int* nout_chosen, bool anti = false);
hipError_t launch_path(int model, int k_assets, int ns, int nout, bool debug, int n_scen, const PathArgs& a,
                       int n_blocks, hipStream_t s, bool anti = false);
// Pathwise Greeks.  Supported programs: one flow date (the last grid point),
// arithmetic ops, and running sums / extremes fed directly by a spot (see
// classify_adjoint).  Outputs per path: price, then d/d params:
//   GBM {s0, sigma, mu}; Heston {s0, v0, kappa, theta, sigma_v, rho};
//   basket {s0_0..s0_{K-1}, sigma_0.., mu_0..}.
// Heston checkpoint interval (grid points).  16 by default: at 32 the
// recomputed segment (variance and two normals per point in registers) pushes
// the kernel past 128 VGPRs and halves its occupancy; 32 is also built.
constexpr int kAdjSeg = 16;
constexpr int kAdjMaxAcc = 4;
constexpr int kAdjMaxOut = 1 + 3 * kMaxAssets;
struct AdjArgs {
  PathArgs p;               // program, model tables, RNG (n_grid points, last = flow date)
  const float* adj_tab;     // [n_grid][2] {sqrt(dt), t log2(e)} (GBM and basket; sqrt(dt) zero at point 0)
  const uint32_t* feed;     // [n_grid] bit q: accumulator q is fed at this point
  const float* rev;         // Heston: [n_grid][8] {t1, t2, t3, t5, dt, sqrt(dt), kappa dt, 0} for the reverse sweep
  float* ck;                // Heston: [n_ck][2][ck_stride] {log2(S/s0), v} before point seg*c
  uint64_t ck_stride;       // >= total paths (checkpoints are indexed by global path)
  int n_acc;
  int acc_slot[kAdjMaxAcc];   // register of the accumulator
  int acc_kind[kAdjMaxAcc];   // OP_AVG, OP_RMAX or OP_RMIN
  int acc_asset[kAdjMaxAcc];  // asset whose spot feeds it
  float sigma[kMaxAssets], mu[kMaxAssets];
  float kappa, theta, sigma_v;
  int seg;  // Heston checkpoint interval: 16 or 32 points
  int tangent;  // Heston: 1 runs the forward (tangent) mode instead of the reverse sweep
};
int adjoint_outputs(int model, int k_assets);
bool adjoint_variant_exists(int model, int k_assets, int n_slots, int* ns_chosen);
hipError_t launch_adjoint(int model, int k_assets, int ns, const AdjArgs& a, int n_blocks, hipStream_t s);

// Pathwise Greeks of a portfolio: one persistent launch per variant over
// (job, block) items of that variant's jobs (16 registers, checkpoint
// interval kAdjSeg).
enum AdjPortfolioVariant {
  APV_GBM = 0,
  APV_HESTON = 1,
  APV_HESTON_FEEDS = 2,
  APV_BASKET5 = 3,
  APV_HESTON_T = 4,        // Heston, forward (tangent) mode
  APV_HESTON_FEEDS_T = 5,  // Heston with running accumulators, tangent mode
  APV_COUNT = 6
};
struct AdjPortfolioArgs {
  const AdjArgs* jobs;    // every adjoint job of the portfolio
  const uint32_t* items;  // [n_items][2] {job, block within the job}
  uint32_t n_items;
  int chunk;              // items taken per queue fetch
  unsigned* counter;      // queue head, zero at launch
  float* ck;              // Heston: checkpoint slot of each workgroup, ck_floats each
  uint64_t ck_floats;     // >= n_ck * 2 * kBlock of every Heston job of the launch
};
// -1 when no portfolio variant covers the job (it runs per contract).
int portfolio_adjoint_variant(int model, int k_assets, int n_slots, int n_acc, bool tangent);
// waves: the register-budget tile of the persistent Greeks kernels, 0 for
// each variant's default budget or 5 (at most 102 VGPRs: 5 waves per SIMD).
int portfolio_adjoint_occupancy(int variant, int waves = 0);
hipError_t launch_portfolio_adjoint(int variant, const AdjPortfolioArgs& q, int grid, hipStream_t s, int waves = 0);
// out[job][kAdjMaxOut]: the first n_out[job] outputs of each job.
hipError_t launch_portfolio_adjoint_finalize(const AdjArgs* jobs, const int* n_blocks, const int* n_out, int n_jobs,
                                             KStats* out, hipStream_t s);

// Longstaff-Schwartz kernels.
constexpr int kLsmBasisMax = 6;
constexpr int kLsmRegCount = (2 * kLsmBasisMax - 1) + kLsmBasisMax + 1;
constexpr int kLsmPassBlocks = 1024;  // grid of the backward passes (grid-stride over paths)
struct LsmPassArgs {
  const float2* cols;
  double* value;
  const double* coef;
  double* reg;
  KStats* stats;
  uint64_t n;
  int n_dates, m, d_b, d_a, final_pass;
  double disc_b, disc_a, disc0;
};
// An exercised payoff that is, at every exercise date, the program
// scale * max(S - strike, floor) (kind 1) or scale * max(strike - S, floor)
// (kind 2) of asset 0 (lsm_vanilla_payoff); kind 0: interpret the program.
struct LsmPayoff {
  int kind;
  float strike, floor, scale;
};
// Writes cols[(k-1)*stride + local path] = {payoff, spot} for k = 1..n_grid-1.
hipError_t launch_lsm_paths(int model, int k_assets, int ns, const PathArgs& a, const LsmPayoff& pay, int n_blocks,
                            float2* cols, uint64_t stride, hipStream_t s);
// reg holds kLsmRegCount x n_blocks doubles; m in [2, kLsmBasisMax].
hipError_t launch_lsm_pass(const LsmPassArgs& p, int n_blocks, hipStream_t s);
// One block: reduce reg and write coef[0..kLsmBasisMax] (coefficients, fitted flag).
hipError_t launch_lsm_solve(const double* reg, int n_blocks, int m, double* coef, hipStream_t s);

// Persistent portfolio launch.  Each job is one contract priced as kt_price
// would (single output, 16-register program); variant picks the path code.
enum PortfolioVariant { PV_GBM = 0, PV_HESTON = 1, PV_BASKET2 = 2, PV_BASKET3 = 3, PV_BASKET4 = 4, PV_BASKET5 = 5 };
constexpr int kPortfolioMaxAssets = 5;
struct PortfolioJob {
  PathArgs a;  // partials: this job's own [n_blocks] range
  int variant;
  int pad;
};
struct PortfolioArgs {
  const PortfolioJob* jobs;
  const uint32_t* items;  // [n_items][2] {job, block within the job}
  uint32_t n_items;
  int chunk;              // items taken per queue fetch
  unsigned* counter;      // queue head (reset by the launch)
};
// Reset the queue and the jobs' error words err[n_jobs], run the persistent
// grid, finalize each job into out[job] and copy its error word to
// err_copy[job].
// The tile of the persistent price kernel (compiled variants): waves, the
// register budget (0 for the compiler's choice, 6 or 8 waves per SIMD: at
// most 80 / 64 VGPRs); unroll, 4-point groups per iteration of the uniform
// runs (1 or 2); lanes per workgroup (256, or 128 with two paths per lane).
struct PortfolioTile {
  int waves = 0, unroll = 1, lanes = kBlock;
};
bool portfolio_tile_valid(const PortfolioTile& t);
hipError_t launch_portfolio(const PortfolioArgs& q, const int* n_blocks, int n_jobs, int grid, unsigned* err,
                            KStats* out, unsigned* err_copy, hipStream_t s, const PortfolioTile& t);
// Resident workgroups (t.lanes lanes each) of the persistent kernel per
// compute unit.
int portfolio_occupancy_blocks_per_cu(const PortfolioTile& t);

// partials[block] = statistics of (a[i] - b[i]) * scale over the block's paths.
hipError_t launch_diff_stats(const double* a, const double* b, double scale, uint64_t n, KStats* partials,
                             hipStream_t s);
// err (nullable): the launch's error word, copied to *err_copy and cleared.
hipError_t launch_finalize(const KStats* partials, int n_blocks, int nout, int n_used, KStats* out,
                           hipStream_t s, unsigned* err = nullptr, unsigned* err_copy = nullptr);
// Greek surface: cells[c] (device) is the seven-scenario launch of cell c
// (ns = 8 or 16 registers), its partials [n_blocks][kScenarioStats] at
// partials + c * n_blocks * kScenarioStats; out[c * kScenarioStats + j].
// n_cu, per_cu: compute units and resident blocks per CU, for the
// co-residency-aware workgroup map (never changes results).
// serial: every cell's serial_tail is set (the scenario-serial kernel).
hipError_t launch_surface(int ns, bool serial, const PathArgs* cells, int n_cells, int n_blocks,
                          const KStats* partials, KStats* out, int n_cu, int per_cu, hipStream_t s);
int surface_occupancy_blocks_per_cu(int ns, bool serial);
hipError_t launch_debug_philox(const uint32_t* ctr, uint32_t k0, uint32_t k1, uint64_t n, uint32_t* out,
                               hipStream_t s);
hipError_t launch_debug_copy(const void* src, void* dst, uint64_t bytes, hipStream_t s);
constexpr int kValuProbeBlocks = 8192;
hipError_t launch_debug_valu(float* out, int iters, hipStream_t s);
// kValuProbeBlocks x 256 lanes, `iters` Philox blocks each (normals: their
// four inverse normals too); out holds one float per lane.
hipError_t launch_debug_rng_rate(bool normals, float* out, int iters, hipStream_t s);
hipError_t launch_debug_normals(const uint32_t* words, uint64_t n, float* out, hipStream_t s);

}  // namespace kt
