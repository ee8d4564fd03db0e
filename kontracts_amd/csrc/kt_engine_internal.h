// kt_engine_internal.h — engine state and helpers shared by the C ABI
// translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <cstring>
#include <memory>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "../../include/kontracts.h"
#include "kt_compiler.h"
#include "kt_launch.h"

struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
};

// Page-locked host buffer: copies from / to it are plain DMA, asynchronous
// to the host, instead of a staged copy through a driver bounce buffer.
struct PinBuf {
  char* p = nullptr;
  size_t cap = 0;
};

// A page-locked staging buffer and the event that marks the last copy out
// of it (waited on before the buffer is rewritten).
struct Staging {
  PinBuf buf;
  hipEvent_t ev = nullptr;
  bool busy = false;
};

struct kt_engine {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  DevBuf arena, partials, stats, errw, pv, z, scratch, paths, aux, book;
  // A second set for a price-only book launched while the Greeks launches of
  // the same call still run (slot 1 of launch_price_book).
  DevBuf book2, partials2, stats2, errw2;
  // persistent portfolio launch: tile choice and the captured launch graph
  int pf_blocks_per_cu = 0;  // 0: the kernel's occupancy limit
  int pf_chunk = 1;
  kt::PortfolioTile pf_tile;  // compiled tile of the price kernel (register budget, unroll, lanes)
  int pf_adj_waves = 0;      // ... of the Greeks kernels: 0 (per-variant default) or 5
  // Direct launches by default: on this stack a hipGraphLaunch of the
  // three-kernel sequence costs more host time than the three launches.
  bool pf_use_graph = false;
  hipGraph_t pf_graph = nullptr;
  hipGraphExec_t pf_exec = nullptr;
  std::vector<uint64_t> pf_key;
  // Longstaff-Schwartz: the ~100 dependent launches of one pricing (paths
  // kernel, backward passes, solves, finalize) captured in a hipGraph and
  // replayed while every launch argument stays the same (lsm_graph option).
  bool lsm_use_graph = true;
  hipGraph_t lsm_graph = nullptr;
  hipGraphExec_t lsm_exec = nullptr;
  std::string lsm_key;
  // Host -> device staging (tables, arguments) and device -> host readback
  // (statistics, error words), both page-locked, in two slots: slot 1 serves
  // a price-only book prepared while slot 0's launches run.
  Staging staging[2];
  PinBuf readback_buf[2];
  // Error word of the per-contract path launches: the finalize kernel copies
  // it next to the statistics and clears it, so it is known zero between
  // calls unless a call was interrupted.
  DevBuf errp;
  bool errp_zero = false;
  // cached occupancies (workgroups per CU) of the price kernel's tiles:
  // [waves 0 / 6 / 8][unroll 1 / 2][lanes 256 / 128]
  int occupancy_portfolio[3][2][2] = {};
  int occupancy_surface[4] = {0, 0, 0, 0};  // Greek surface kernel: 8 / 16 registers, then their serial-tail variants
  int occupancy_adjoint[2][kt::APV_COUNT] = {};  // Greeks kernels at the default budget / 5 waves
  bool pathwise_tangent = false;  // Heston pathwise Greeks: forward tangents instead of the reverse sweep
  bool profiling = false;
  // Profiling brackets: event pairs recorded on the stream and resolved when
  // the totals are read (no host wait inside a call).
  struct ProfBracket {
    hipEvent_t a, b;
    double path_steps;
  };
  std::vector<hipEvent_t> prof_pool;
  std::vector<ProfBracket> prof_open;
  hipEvent_t prof_cur = nullptr;
  double prof_ms = 0.0, prof_path_steps = 0.0;
  uint64_t prof_launches = 0;
  void prof_resolve();
  int adj_seg = kt::kAdjSeg;  // Heston adjoint checkpoint interval (kt_engine_set_option)
  // Lowered programs of recent (contract, grid, asset names): a book priced
  // again (new market data, same contracts) skips the lowering.  Keyed by
  // the parse cache's contract (its address, held by the entry), the asset
  // names and a hash of the grid; an entry keeps its grid, compared on every
  // hit.
  struct Lowered {
    std::vector<double> grid;
    std::shared_ptr<const kt::Program> prog;
    std::shared_ptr<const kt::Con> con;
  };
  std::unordered_map<std::string, Lowered> lower_cache;
  // classify_adjoint of cached programs (null plan: not pathwise-Greek-able),
  // cleared with lower_cache.
  std::unordered_map<const kt::Program*, std::shared_ptr<const kt::AdjPlan>> adj_plan_cache;
  // Parsed contracts and their plans of recent contract JSON texts.
  struct Parsed {
    std::shared_ptr<const kt::Con> con;
    kt::Plan plan;
  };
  std::unordered_map<std::string, Parsed> parse_cache;
  // Bracket a group of launches with events when profiling is enabled.
  void prof_begin();
  void prof_end(double path_steps);
  ~kt_engine();
  void* ensure(DevBuf& b, size_t bytes);
  std::vector<void*> upload(const std::vector<std::pair<const void*, size_t>>& blobs);
  // Staging buffer of at least `bytes`, free to write (an earlier copy out of
  // it has completed); stage_copy sends bytes [offset, offset + bytes) of it
  // to the same range of dev.
  char* stage(size_t bytes, int slot = 0);
  void stage_copy(void* dev, size_t bytes, int slot = 0, size_t offset = 0);
  // Readback buffer of at least `bytes` (valid after the stream syncs).
  char* readback(size_t bytes, int slot = 0);
};

namespace kt {

struct ModelDesc {
  int kind = 0;
  std::vector<std::string> names;
  std::vector<double> s0, mu, sigma;
  std::vector<double> chol;  // K x K lower-triangular (correlated GBM)
  double v0 = 0, kappa = 0, theta = 0, sigma_v = 0, rho = 0;
  int scheme = 0;
};

struct PreparedContract {
  std::shared_ptr<const Con> con;
  std::vector<double> grid;
};

extern const double kZ95;
extern const double kLog2e;

// 1 when rows 1 .. n_rows - 1 of a per-point table (stride floats a row) are
// bitwise identical: a grid of equal steps, whose coefficients the path
// kernels then hold in registers.
inline int uniform_table_rows(const std::vector<float>& tab, size_t n_rows, size_t stride) {
  if (n_rows < 3 || tab.size() < n_rows * stride) return 0;
  for (size_t k = 2; k < n_rows; ++k)
    if (std::memcmp(&tab[k * stride], &tab[stride], stride * sizeof(float)) != 0) return 0;
  return 1;
}

ModelDesc model_from_abi(const kt_model* m);
void validate_model(const ModelDesc& d);
int model_normals(const ModelDesc& d);
int model_kernel_kind(const ModelDesc& d);
std::vector<float> model_table(const ModelDesc& d, const std::vector<double>& grid, int* stride);
kt_price_result result_from_stats(const KStats& s);
void check_hip(hipError_t e, const char* what);
int set_error(kt_engine* E, int code, const std::string& msg);
int lowest_error_code(unsigned bits);
const char* error_text(int code);
// error_text with the program's payload: TimeOutOfRange names the fixing
// date ("...: <date>"), as KontractError::TimeOutOfRange(f64) carries it.
std::string error_text(int code, const Program& prog);

// lower({con}, grid, names), memoised on the engine (the program depends on
// nothing else); con comes from prepare() with the engine, so one contract
// text maps to one con while its parse-cache entry lives.
std::shared_ptr<const Program> lower_cached(kt_engine* E, const std::shared_ptr<const Con>& con,
                                            const std::vector<double>& grid, const std::vector<std::string>& names);
// classify_adjoint of a program returned by lower_cached, memoised with it;
// null when the program has no pathwise Greeks (Unsupported).
std::shared_ptr<const AdjPlan> adjoint_plan_cached(kt_engine* E, const std::shared_ptr<const Program>& prog);

// allow_vr: the caller handles cfg's variance-reduction flags; otherwise
// they are rejected.
// E (optional): parse through the engine's cache of contract texts.
PreparedContract prepare(const char* json, const ModelDesc& md, const kt_mc_config* cfg, bool allow_vr = false,
                         kt_engine* E = nullptr);
// pv_keep (device, n_paths doubles, nullable): receives each path's pv and
// stays on the device.
std::vector<KStats> run_program(kt_engine* E, const Program& prog, const ModelDesc& md,
                                const std::vector<double>& grid, double rate, uint64_t n_paths, uint64_t seed,
                                double* pv_host, float* z_host, double* pv_keep = nullptr);

std::vector<KGroup> knockout_tables(const Program& prog, const std::vector<ModelDesc>& models,
                                    std::vector<float>* thr);
// exp(-rate t_k) at the grid points where the program pays a flow, NaN at
// the others (no kernel reads them); out[0 .. grid.size()).
void discount_factors(const Program& prog, const std::vector<double>& grid, double rate, double* out);

// Model-dependent tables of one launch of the path kernel: per-point model
// table (scenario-interleaved when SC > 1), discount tables per distinct
// rate, and the knock-out groups with their thresholds.
struct LaunchTables {
  std::vector<float> tab;
  int stride = 0;
  std::vector<double> df;
  std::vector<int> df_index;
  std::vector<KGroup> groups;
  std::vector<float> thr;
};
// discount = false leaves df empty (a caller that shares one discount table).
LaunchTables launch_tables(const Program& prog, const std::vector<ModelDesc>& models,
                           const std::vector<double>& rates, const std::vector<double>& grid, bool discount = true);

// Model constants of the path kernels (spots, Cholesky factor, Heston
// correlation split) and the Philox key.
void fill_model_args(PathArgs& A, const ModelDesc& md, uint64_t seed);

// The seven device tables of one path-kernel launch, in the order
// program_path_args expects their device copies (ops receives the op list,
// padded to one no-op when empty, and must outlive the blobs).
std::vector<std::pair<const void*, size_t>> program_blobs(const Program& prog, const LaunchTables& T,
                                                         std::vector<KOp>* ops);
// Kernel arguments of one launch (scenario sc of models[sc]) whose tables
// sit at dev[0..6]; partials, error word and path range are left to the
// caller.
PathArgs program_path_args(const Program& prog, const std::vector<ModelDesc>& models, const LaunchTables& T,
                           const std::vector<double>& grid, uint64_t seed, const std::vector<void*>& dev);

std::vector<KStats> run_program_scen(kt_engine* E, const Program& prog, const std::vector<ModelDesc>& models,
                                     const std::vector<double>& rates, const std::vector<std::vector<double>>& stat_w,
                                     const std::vector<double>& grid, uint64_t n_paths, uint64_t seed,
                                     double* pv_host, float* z_host, double* pv_keep = nullptr, bool anti = false,
                                     int combine = 0);

// Implemented in kt_greeks.cpp / kt_lsm.cpp / kt_portfolio.cpp.
void greeks_gbm_crn(kt_engine* E, const char* json, const char* asset, double s0, double sigma,
                    const kt_mc_config* cfg, const kt_bumps& h, kt_greeks* out);
void greek_surface(kt_engine* E, const char* json, const char* asset, const double* spots, uint64_t n_spots,
                   const double* vols, uint64_t n_vols, const kt_mc_config* cfg, const kt_bumps& h, kt_greeks* out);
void greeks_bump(kt_engine* E, const char* json, const kt_model* model, const kt_mc_config* cfg, const double* h,
                 uint64_t n_params, double* out, double* stderr_out);
// stats_out (nullable): the per-output statistics (price first).
void greeks_adjoint(kt_engine* E, const char* json, const kt_model* model, const kt_mc_config* cfg, double* out,
                    double* stderr_out, uint64_t n_out, std::vector<KStats>* stats_out = nullptr);
// pv_host[n_paths] and coef_host[(n_dates-1) x (n_basis+1)] (nullable)
// receive each path's discounted value and each date's regression
// coefficients followed by a fitted flag.
void price_american(kt_engine* E, const char* json, const double* dates, uint64_t n_dates, const kt_model* model,
                    const kt_mc_config* cfg, int n_basis, kt_price_result* out, double* pv_host = nullptr,
                    double* coef_host = nullptr);
// job_index (nullable): the caller's number of each job, for messages.
void price_portfolio(kt_engine* E, const kt_job* jobs, uint64_t n_jobs, kt_price_result* out,
                     const uint64_t* job_index = nullptr);
// A price-only book in two halves: launch_price_book prepares the jobs,
// uploads through staging slot `slot` into that slot's device buffers and
// enqueues the persistent launch and the statistics' readback without
// waiting; finish_price_book waits for the stream and writes out[] (or
// raises the first failing job's error, as a loop of price_gbm would).  The
// Greeks book launches its price-only jobs in slot 1 before it prepares its
// own launches in slot 0.
struct PriceBookRun;
struct PriceBookRunDeleter {
  void operator()(PriceBookRun* r) const;
};
using PriceBookHandle = std::unique_ptr<PriceBookRun, PriceBookRunDeleter>;
PriceBookHandle launch_price_book(kt_engine* E, const kt_job* jobs, uint64_t n_jobs, kt_price_result* out,
                                  const uint64_t* job_index, int slot);
void finish_price_book(kt_engine* E, PriceBookRun& run);
// Host tables of one pathwise-Greeks job (the per-contract launch and the
// portfolio share them): blobs in the order fill_adjoint_args expects.
struct AdjJobHost {
  std::vector<KOp> ops;
  std::vector<float> tab, adj, thr, rev;
  std::vector<double> df;
  int stride = 0;
  AdjPlan plan;
};
AdjJobHost adjoint_job_host(const Program& prog, const ModelDesc& md, const std::vector<double>& grid, double rate,
                            AdjPlan plan);
std::vector<std::pair<const void*, size_t>> adjoint_blobs(const AdjJobHost& H, const Program& prog);
void fill_adjoint_args(AdjArgs& A, const AdjJobHost& H, const Program& prog, const ModelDesc& md,
                       const std::vector<double>& grid, uint64_t seed, const std::vector<void*>& dev);
// Pricing of a portfolio with the pathwise Greeks of every job they cover.
void price_portfolio_greeks(kt_engine* E, const kt_job* jobs, uint64_t n_jobs, kt_price_result* out, double* greeks,
                            double* greeks_se, uint64_t stride);

// kt_vr.cpp: price_gbm with variance reduction (antithetic pairs and/or the
// at-the-money call control variate).
kt_price_result price_vr(kt_engine* E, const Con& con, const ModelDesc& md, const std::vector<double>& grid,
                         double rate, uint64_t n_paths, uint64_t seed, bool antithetic, bool control_variate);
// The reference's Black-Scholes call (Abramowitz-Stegun normal CDF).
double bs_call_reference(double s, double k, double t, double r, double sigma);

// Scratch array of the allocation-free entry point on `device` (set up
// once per device by kt_engine_create, with the device current).
hipError_t noalloc_scratch_init(int device);

}  // namespace kt

// C ABI guard: kt::Error, allocation failures and other exceptions become a
// kt_status and the engine's (and the thread's) last error message.
#define KT_GUARD_BEGIN try {
#define KT_GUARD_END(E)                                                                   \
  }                                                                                       \
  catch (const kt::Error& e) {                                                            \
    return (kt_status)kt::set_error((E), e.code, e.msg);                                  \
  }                                                                                       \
  catch (const std::bad_alloc&) {                                                         \
    return (kt_status)kt::set_error((E), KT_ERR_OUT_OF_MEMORY, "host allocation failed"); \
  }                                                                                       \
  catch (const std::exception& ex) {                                                      \
    return (kt_status)kt::set_error((E), KT_ERR_DEVICE, ex.what());                       \
  }

