/*
 * kontracts.h — C ABI of the MI355X Monte Carlo pricing engine.
 *
 * This is the drop-in boundary for kontract's path-simulation hot path.  Every
 * entry point replaces one reference function; the citation on each names the
 * Rust function (file:line in the kontract crate) whose argument meaning,
 * defaults and error behaviour it keeps.
 *
 * Conventions
 *   - Plain C types only: pointers, sizes, doubles.  The caller owns every host
 *     buffer; the engine owns device memory.
 *   - Contracts cross the boundary as serde's externally tagged JSON, exactly as
 *     Contract::to_json writes it (src/ast.rs:210-218), e.g.
 *       {"When":[{"At":1.0},{"Scale":[{"Max":[{"Sub":[{"Spot":"X"},{"Const":100.0}]},
 *        {"Const":0.0}]},{"One":"USD"}]}]}
 *   - kt_status values 1..7 are KontractError's variants (src/error.rs:7-28) in
 *     declaration order; the message is available through kt_last_error.
 *   - Calls are synchronous: each returns after the engine's HIP stream has
 *     drained.  An engine is not thread-safe; use one engine per host thread.
 *   - Results depend only on the inputs and the seed (counter-based RNG, fixed
 *     reduction order), not on launch shape or scheduling.
 */
#ifndef KONTRACTS_H
#define KONTRACTS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KT_ABI_VERSION 1
#define KT_MAX_ASSETS 8

typedef enum kt_status {
  KT_OK = 0,
  KT_ERR_UNKNOWN_ASSET = 1,       /* KontractError::UnknownAsset */
  KT_ERR_TIME_OUT_OF_RANGE = 2,   /* KontractError::TimeOutOfRange */
  KT_ERR_STEP_OUT_OF_RANGE = 3,   /* KontractError::StepOutOfRange */
  KT_ERR_INCONSISTENT_PATH = 4,   /* KontractError::InconsistentPath */
  KT_ERR_SERDE = 5,               /* KontractError::Serde */
  KT_ERR_MALFORMED_CONTRACT = 6,  /* KontractError::MalformedContract */
  KT_ERR_UNSUPPORTED = 7,         /* KontractError::Unsupported */
  KT_ERR_INVALID_ARGUMENT = 100,  /* bad pointer / size at the C boundary */
  KT_ERR_DEVICE = 101,            /* HIP runtime or kernel failure */
  KT_ERR_OUT_OF_MEMORY = 102
} kt_status;

typedef struct kt_engine kt_engine;

/* Model kinds: the three simulators on the hot path. */
enum {
  KT_MODEL_GBM = 0,            /* Gbm (src/simulator.rs:188-253) */
  KT_MODEL_HESTON = 1,         /* HestonSimulator (src/simulator.rs:453-584) */
  KT_MODEL_CORRELATED_GBM = 2  /* CorrelatedGbmN (src/multi_asset.rs:49-184) */
};

/* Heston variance schemes.  FLOOR is the reference's: v <- max(v_next, 0) after
 * every step (src/simulator.rs:531-533).  FULL_TRUNCATION uses max(v, 0) inside
 * the drift and diffusion and keeps v unfloored (Lord, Koekkoek, van Dijk 2010). */
enum { KT_HESTON_FLOOR = 0, KT_HESTON_FULL_TRUNCATION = 1 };

/* A simulator.  Replaces `&dyn Simulator` (src/simulator.rs:70-178).
 *   GBM:            n_assets = 1; s0, mu (= r - q, bindings.rs:518-521), sigma.
 *   HESTON:         n_assets = 1; s0 = spot, mu = r (drift), v0..rho below.
 *   CORRELATED_GBM: n_assets = k (2..8); s0/mu/sigma per factor (GbmFactor,
 *                   src/fx.rs:149-170) and corr = k*k row-major correlation. */
typedef struct kt_model {
  int32_t kind;
  int32_t n_assets;
  const char* const* asset_names; /* n_assets NUL-terminated names */
  const double* s0;               /* [n_assets] */
  const double* mu;               /* [n_assets] */
  const double* sigma;            /* [n_assets] (unused for HESTON) */
  const double* corr;             /* [n_assets*n_assets] (CORRELATED_GBM only) */
  double v0, kappa, theta, sigma_v, rho; /* HESTON only */
  int32_t heston_scheme;          /* KT_HESTON_FLOOR (default) or FULL_TRUNCATION */
  int32_t reserved;
} kt_model;

/* McConfig (src/pricer.rs:30-54).  grid/n_grid, when set, replace the compiled
 * time grid, which is the reference's price_on_paths route on an explicit grid
 * (src/pricer.rs:230-241).  antithetic/control_variate are
 * VarianceReductionConfig (src/variance_reduction.rs:35-43); kt_price,
 * kt_greeks_gbm and kt_price_portfolio apply them as price_gbm does,
 * kt_price_batch ignores them as price_batch_gbm does, and the other entry
 * points return KT_ERR_UNSUPPORTED when they are set. */
typedef struct kt_mc_config {
  uint64_t n_paths;
  uint64_t seed;
  uint64_t steps_per_year;
  double rate;
  const double* grid;
  uint64_t n_grid;
  uint32_t antithetic;
  uint32_t control_variate;
} kt_mc_config;

/* PriceResult (src/pricer.rs:58-71). */
typedef struct kt_price_result {
  double price;
  double sample_std;
  double std_error;
  double ci95_low;
  double ci95_high;
  uint64_t n_paths;
} kt_price_result;

/* BumpSizes (src/greeks.rs:22-39); defaults spot 1e-2, vol 1e-2, rate 1e-4. */
typedef struct kt_bumps {
  double spot;
  double vol;
  double rate;
} kt_bumps;

/* Greeks (src/greeks.rs:43-54) plus the Monte Carlo standard error of each
 * estimate, computed from the per-path finite-difference samples. */
typedef struct kt_greeks {
  double price, delta, gamma, vega, rho;
  double std_error[5]; /* same order: price, delta, gamma, vega, rho */
} kt_greeks;

/* Engine lifetime. `device` is the HIP device ordinal. */
kt_status kt_engine_create(int device, kt_engine** out);
void kt_engine_destroy(kt_engine* engine);
/* Message of the last failing call on this engine ("" if none). */
const char* kt_last_error(const kt_engine* engine);
/* Message of the last failing engine-less call on this thread. */
const char* kt_last_error_global(void);

/* price_gbm (src/pricer.rs:97-115): compile, build the grid
 * (time_grid_forced with the model's fine-grid requirement), simulate, evaluate
 * cashflows per path and summarize.  Replaces PyContract::price
 * (src/bindings.rs:350-379) underneath the Python API.
 * With variance reduction (src/pricer.rs:117-228):
 *  - antithetic, GBM: n_paths / 2 pairs (Z, -Z); the result summarizes the
 *    pair means (n_paths of the result = pairs).  Other models ignore it.
 *  - control_variate: beta = 1 on an at-the-money call
 *    when(at(T), scale(max(S - s0, 0), one(USD))) on the model's first asset,
 *    T = last grid date, corrected by its Black-Scholes value at
 *    (s0, sigma) of a GBM, (100, 0.2) for the other models.
 *  - both (GBM): price = mean_pairs(target) - (mean_pairs(control) - BS),
 *    sample_std of the target's pair means, std_error over 2 x pairs. */
kt_status kt_price(kt_engine* engine, const char* contract_json, const kt_model* model,
                   const kt_mc_config* cfg, kt_price_result* out);

/* kt_price without device allocation.  No counterpart in the reference: it is
 * for callers that cannot call hipMalloc (code under HIP graph capture, tight
 * serving loops, harnesses that forbid allocation).  Same arguments, result
 * and errors as kt_price (bit-identical results), except:
 *  - the program tables, block partials, error word and statistics live in
 *    one static 8 MiB device array per GPU (held until the process ends), so
 *    n_paths is at most 2^26 (67,108,864) and the tables at most 2 MiB less
 *    4 KiB; larger calls return KT_ERR_INVALID_ARGUMENT before any launch;
 *  - variance-reduction flags return KT_ERR_UNSUPPORTED;
 *  - every call on one GPU shares that array and leaves its contents there:
 *    calls that overlap on one GPU (two engines, two host threads) race, and
 *    callers must serialise them. */
kt_status kt_price_noalloc(kt_engine* engine, const char* contract_json, const kt_model* model,
                           const kt_mc_config* cfg, kt_price_result* out);

/* price_batch_gbm (src/pricer.rs:250-294): one merged grid (union of dates,
 * max horizon, OR of fine flags), one simulation shared by all contracts, one
 * result per contract.  n_contracts == 0 returns KT_OK with nothing written. */
kt_status kt_price_batch(kt_engine* engine, const char* const* contract_jsons, uint64_t n_contracts,
                         const kt_model* model, const kt_mc_config* cfg, kt_price_result* out);

/* greeks_gbm (src/greeks.rs:60-98): CRN bump-and-reprice of a one-asset GBM with
 * drift = cfg->rate; the seven repricings share one simulation of the
 * normals on the GPU.  bumps == NULL means BumpSizes::default().  With
 * variance reduction each repricing is a variance-reduced kt_price at the same
 * seed; std_error then holds the price's only (the others are NaN). */
kt_status kt_greeks_gbm(kt_engine* engine, const char* contract_json, const char* asset, double s0,
                        double sigma, const kt_mc_config* cfg, const kt_bumps* bumps, kt_greeks* out);

/* greek_surface (src/surface.rs:104-138): greeks_gbm at every cell of the
 * (spots x vols) grid, out[i * n_vols + j] for (spots[i], vols[j]) (the
 * reference's [spot][vol] matrices; rho and the standard errors ride along).
 * Every cell uses cfg->seed, as the reference's cells do; all cells run as
 * one launch, each cell's result equal to kt_greeks_gbm at that point.  An
 * empty axis returns KT_ERR_INCONSISTENT_PATH; a failing cell fails the
 * call (the lowest cell index wins). */
kt_status kt_greek_surface(kt_engine* engine, const char* contract_json, const char* asset, const double* spots,
                           uint64_t n_spots, const double* vols, uint64_t n_vols, const kt_mc_config* cfg,
                           const kt_bumps* bumps, kt_greeks* out);

/* CRN bump-and-reprice for any model, the greeks_gbm pattern (central
 * differences, common random numbers; src/greeks.rs:69-89) applied to the
 * model's parameters.  Parameters: GBM {s0, sigma, mu}; HESTON {s0, v0, kappa,
 * theta, sigma_v, rho}.  h[i] are absolute bumps (n_params entries).
 * out[0] = price, out[1+i] = dPrice/dparam_i; stderr_out (nullable) holds the
 * standard error of each (for a derivative: of the per-path differences of
 * the two bumped present values).  Reprices with separate launches at the
 * same seed. */
kt_status kt_greeks_bump(kt_engine* engine, const char* contract_json, const kt_model* model,
                         const kt_mc_config* cfg, const double* h, uint64_t n_params, double* out,
                         double* stderr_out);

/* Pathwise (adjoint) Greeks for Lipschitz payoffs paid on one date (the
 * last grid point): arithmetic functions of the spots there and of running
 * averages / extremes of one asset's spot (European, Asian, lookback,
 * basket).  Anything else (barriers, digitals, price-dependent conditions,
 * flows on several dates) returns KT_ERR_UNSUPPORTED: use kt_greeks_bump.
 * Parameters: GBM {s0, sigma, mu}; HESTON {s0, v0, kappa, theta, sigma_v,
 * rho}; CORRELATED_GBM {s0 of each asset, sigma of each, mu of each}.
 * out[0] = price, out[1+i] = dPrice/dparam_i (n_out >= 1 + #params);
 * stderr_out (nullable) holds each one's Monte Carlo standard error.  Heston
 * runs a reverse sweep that recomputes each segment of the path (16 grid
 * points by default, kt_engine_set_option "adjoint_segment") from a
 * checkpoint.  Heston pathwise derivatives assume the variance seldom reaches
 * the floor (Feller: 2 kappa theta >= sigma_v^2); where it often does, sqrt(v)
 * is not Lipschitz and bump Greeks are the unbiased choice.  No reference counterpart (the reference has bump Greeks
 * only, src/greeks.rs). */
kt_status kt_greeks_adjoint(kt_engine* engine, const char* contract_json, const kt_model* model,
                            const kt_mc_config* cfg, double* out, double* stderr_out, uint64_t n_out);

/* price_american_lsm (src/lsm.rs:80-124): Longstaff-Schwartz on the grid
 * [0, dates...], ITM regression on {1, S, .., S^(n_basis-1)} with normal
 * equations accumulated and solved in f64 (poly_fit/solve_linear,
 * src/lsm.rs:259-288, src/numerics.rs:190-236).  The exercised payoff must
 * have all its flows at the exercise date (no when/until/anytime inside). */
kt_status kt_price_american(kt_engine* engine, const char* payoff_json, const double* exercise_dates,
                            uint64_t n_dates, const kt_model* model, const kt_mc_config* cfg, int32_t n_basis,
                            kt_price_result* out);

/* One contract of a heterogeneous portfolio: its own model and MC config.
 * Semantically a loop of price_gbm calls (src/pricer.rs:97-115). */
typedef struct kt_job {
  const char* contract_json;
  kt_model model;
  kt_mc_config cfg;
} kt_job;

/* Prices n_jobs contracts with different models, grids and path counts in one
 * persistent launch over a (contract, path-block) work queue.  Each result is
 * identical to kt_price on that job alone. */
kt_status kt_price_portfolio(kt_engine* engine, const kt_job* jobs, uint64_t n_jobs, kt_price_result* out);

/* kt_price_portfolio plus the pathwise Greeks of every job that admits them
 * (the kt_greeks_adjoint payoffs: one flow date, arithmetic ops, running
 * sums / extremes fed by a spot).  Row j of greeks / greeks_se (stride >= 24
 * doubles, nullable) holds the derivatives of job j in kt_greeks_adjoint's
 * order without the price (GBM {s0, sigma, mu}; Heston {s0, v0, kappa, theta,
 * sigma_v, rho}; basket {s0_a, sigma_a, mu_a}), NaN for jobs priced without
 * Greeks (barriers, digitals, variance-reduced configurations, programs too
 * large for the Greeks kernels).  For a job with Greeks, out[j] and its row
 * equal kt_greeks_adjoint on that job alone (its price comes from the Greeks
 * kernel); for the others out[j] equals kt_price.  Covered jobs run in
 * persistent launches over (job, block) queues, one per kernel variant. */
kt_status kt_price_portfolio_greeks(kt_engine* engine, const kt_job* jobs, uint64_t n_jobs, kt_price_result* out,
                                    double* greeks, double* greeks_se, uint64_t stride);

/* Kernel timing: when enabled, HIP events on the engine's stream bracket every
 * path-kernel launch group; the totals accumulate until reset.  path_steps
 * counts paths x grid intervals x scenarios simulated by those launches. */
kt_status kt_engine_set_profiling(kt_engine* engine, int32_t enabled);
kt_status kt_engine_kernel_time(kt_engine* engine, double* total_ms, uint64_t* n_launches, double* path_steps,
                                int32_t reset);

/* Tuning knobs of an engine (no reference counterpart).  Names:
 *   "adjoint_segment": Heston pathwise-Greek checkpoint interval, 16 or 32
 *                      grid points (default 16).
 *   "portfolio_blocks_per_cu": resident blocks per compute unit of the
 *                      persistent portfolio grid (0 = the kernel's occupancy
 *                      limit, the default).
 *   "portfolio_chunk": work items taken per queue fetch (default 1).
 *   "pathwise_mode":   Heston pathwise Greeks by the reverse (adjoint) sweep
 *                      with checkpoints, 0 (default), or by forward tangents
 *                      of the five variance parameters in one pass, 1 (same
 *                      estimator to rounding; no checkpoint memory).
 *   "portfolio_graph": 1 replays the portfolio's launch sequence from a
 *                      captured hipGraph while the book keeps its shape; 0
 *                      (default) launches the three kernels directly, which
 *                      measured faster on ROCm for books of 10-100 jobs.
 *   "portfolio_waves": register budget of the persistent price kernel: 0
 *                      (default: the compiler's), 6 or 8 waves per SIMD.
 *   "portfolio_unroll": 4-point groups per iteration of the persistent price
 *                      kernel's uniform-run loops: 1 (default) or 2.
 *   "portfolio_lanes": lanes per workgroup of the persistent price kernel:
 *                      256 (default, one path per lane) or 128 (each lane
 *                      prices two paths of a 256-path block in turn).
 *   "portfolio_adjoint_waves": register budget of the persistent Greeks
 *                      kernels: 0 (default: 102 VGPRs for GBM and Heston
 *                      with running accumulators, 85 for baskets, 128 for
 *                      Heston) or 5 waves per SIMD for all.
 *   "lsm_graph":       1 (default) replays kt_price_american's launch chain
 *                      (paths kernel, backward passes, solves) from a
 *                      captured hipGraph while its launch arguments are
 *                      unchanged; 0 launches each kernel directly.
 * Results never depend on these: each job's reduction order is fixed.
 * Unknown names or values return KT_ERR_INVALID_ARGUMENT. */
kt_status kt_engine_set_option(kt_engine* engine, const char* name, int64_t value);

/* ---- Host-only (no GPU needed) -------------------------------------------
 * compile() + Plan::time_grid_forced (src/compiler.rs:54-85).  Writes up to
 * cap grid points and dates; *n_grid / *n_dates receive the full counts.
 * assets receives the sorted asset names joined by '\n'. */
kt_status kt_plan(const char* contract_json, uint64_t steps_per_year, int32_t force_fine, double* grid,
                  uint64_t* n_grid, double* dates, uint64_t* n_dates, uint64_t cap, int32_t* needs_fine_grid,
                  double* horizon, char* assets, uint64_t assets_cap);

/* black_scholes_call (src/numerics.rs:63-71) with the reference's
 * Abramowitz-Stegun 7.1.26 normal CDF (src/numerics.rs:22-41): the value the
 * control variate corrects with.  Host only. */
double kt_black_scholes_call(double s, double k, double t, double r, double sigma);

/* Text listing of the per-step op program the engine would run for a
 * contract on a grid (diagnostics; host only). */
kt_status kt_describe_program(const char* contract_json, const kt_model* model, const kt_mc_config* cfg,
                              char* buf, uint64_t cap);

/* ---- Parity hooks (used by the tests) ---------------------------------------
 * Philox4x32-10 on the GPU: n counters (4 words each) under one key. */
kt_status kt_debug_philox(kt_engine* engine, const uint32_t* ctr, const uint32_t* key, uint64_t n,
                          uint32_t* out);
/* FP32 inverse normal on the GPU applied to raw 32-bit words. */
kt_status kt_debug_normals(kt_engine* engine, const uint32_t* words, uint64_t n, float* out);
/* Streaming device-to-device copy of `bytes` (multiple of 16) through a plain
 * copy kernel, repeated `reps` times after one warm launch; *ms receives the
 * average kernel time.
 * Calibrates the profiler's HBM byte counters against a known size. */
kt_status kt_debug_copy(kt_engine* engine, uint64_t bytes, int32_t reps, double* ms);
/* VALU issue calibration: 2^21 lanes each run `iters` (a multiple of 8)
 * iterations of eight independent FP32 FMAs (eight VALU instructions per
 * iteration, the loop unrolled by 8), repeated
 * `reps` times after one warm launch; *ms receives the average kernel time
 * and *lane_ops the FMA lane-ops per second.  Measures the vector issue rate
 * the path kernels' roofline is compared with. */
kt_status kt_debug_valu(kt_engine* engine, int32_t iters, int32_t reps, double* ms, double* lane_ops);
/* Normal-generator throughput: 2^21 lanes each draw `iters` Philox4x32-10
 * blocks with the path kernels' counter layout and fold the four words
 * (normals == 0) or their four FP32 inverse normals (normals != 0) into one
 * accumulator, repeated `reps` times after one warm launch; *ms receives the
 * average kernel time and *words the words (or normals) per second.  Gives
 * the issue cost of the generator the path kernels run once per normal. */
kt_status kt_debug_rng_rate(kt_engine* engine, int32_t normals, int32_t iters, int32_t reps, double* ms,
                            double* words);
/* kt_price that also returns every path's present value (pv_out[n_paths]) and,
 * when normals_out != NULL, the FP32 normals the GPU drew, laid out
 * [path][grid point][component].  grid_out receives the grid (capacity
 * grid_cap) and *n_grid its length. */
kt_status kt_debug_paths(kt_engine* engine, const char* contract_json, const kt_model* model,
                         const kt_mc_config* cfg, double* pv_out, float* normals_out, double* grid_out,
                         uint64_t grid_cap, uint64_t* n_grid, kt_price_result* out);
/* kt_price_american that also returns every path's value discounted to t = 0
 * (pv_out[n_paths]) and, when coef_out != NULL, each regression date's fitted
 * coefficients: row d (d = 0 .. n_dates-2) holds c_0 .. c_(n_basis-1) and then
 * 1.0 if a regression was fitted (>= n_basis ITM paths), else 0.0. */
kt_status kt_debug_american(kt_engine* engine, const char* payoff_json, const double* exercise_dates,
                            uint64_t n_dates, const kt_model* model, const kt_mc_config* cfg, int32_t n_basis,
                            double* pv_out, double* coef_out, kt_price_result* out);

#ifdef __cplusplus
}
#endif
#endif /* KONTRACTS_H */
