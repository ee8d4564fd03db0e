// kt_engine.cpp — the C ABI (include/kontracts.h): validation, model tables,
// device buffers, kernel launches and result assembly.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <limits>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/kontracts.h"
#include "kt_compiler.h"
#include "kt_engine_internal.h"
#include "kt_launch.h"

using namespace kt;

namespace kt {

thread_local std::string g_last_error;

const double kZ95 = 1.959963984540054;  // src/pricer.rs:74
const double kLog2e = 1.4426950408889634;

// ---------------------------------------------------------------------------
// Models
// ---------------------------------------------------------------------------
ModelDesc model_from_abi(const kt_model* m) {
  if (!m) raise(KT_ERR_INVALID_ARGUMENT, "null model");
  ModelDesc d;
  d.kind = m->kind;
  int k = m->n_assets;
  if (m->kind == KT_MODEL_GBM || m->kind == KT_MODEL_HESTON) {
    if (k != 1) raise(KT_ERR_INVALID_ARGUMENT, "GBM and Heston models have exactly one asset");
  } else if (m->kind == KT_MODEL_CORRELATED_GBM) {
    // CorrelatedGbmN::new (src/multi_asset.rs:62-83)
    if (k < 2) raise(KT_ERR_MALFORMED_CONTRACT, "CorrelatedGbmN: at least 2 assets required");
    if (k > KT_MAX_ASSETS) raise(KT_ERR_UNSUPPORTED, "CorrelatedGbmN: at most 8 assets on this engine");
    if (!m->corr) raise(KT_ERR_MALFORMED_CONTRACT, "CorrelatedGbmN: correlation matrix must be kxk");
  } else {
    raise(KT_ERR_INVALID_ARGUMENT, "unknown model kind");
  }
  if (!m->asset_names || !m->s0 || !m->mu) raise(KT_ERR_INVALID_ARGUMENT, "null model arrays");
  if (m->kind != KT_MODEL_HESTON && !m->sigma) raise(KT_ERR_INVALID_ARGUMENT, "null sigma");
  for (int i = 0; i < k; ++i) {
    if (!m->asset_names[i]) raise(KT_ERR_INVALID_ARGUMENT, "null asset name");
    d.names.push_back(m->asset_names[i]);
    d.s0.push_back(m->s0[i]);
    d.mu.push_back(m->mu[i]);
    d.sigma.push_back(m->kind == KT_MODEL_HESTON ? 0.0 : m->sigma[i]);
  }
  if (m->kind == KT_MODEL_CORRELATED_GBM) {
    for (int i = 0; i < k; ++i)
      if (d.sigma[i] < 0.0) raise(KT_ERR_MALFORMED_CONTRACT, "CorrelatedGbmN: all volatilities must be >= 0");
    // cholesky_lower with the PSD floor (src/numerics.rs:158-183)
    d.chol.assign((size_t)k * k, 0.0);
    for (int i = 0; i < k; ++i)
      for (int j = 0; j <= i; ++j) {
        double sum = m->corr[i * k + j];
        for (int q = 0; q < j; ++q) sum -= d.chol[i * k + q] * d.chol[j * k + q];
        if (i == j) {
          d.chol[i * k + j] = std::sqrt(std::max(sum, 0.0));
        } else {
          double piv = d.chol[j * k + j];
          d.chol[i * k + j] = piv > 0.0 ? sum / piv : 0.0;
        }
      }
  }
  d.v0 = m->v0;
  d.kappa = m->kappa;
  d.theta = m->theta;
  d.sigma_v = m->sigma_v;
  d.rho = m->rho;
  d.scheme = m->heston_scheme;
  return d;
}

// Parameter checks done by the reference simulators before simulating
// (Gbm::simulate src/simulator.rs:221-225, HestonSimulator::simulate :556-580).
void validate_model(const ModelDesc& d) {
  if (d.kind == KT_MODEL_GBM && d.sigma[0] < 0.0) raise(KT_ERR_MALFORMED_CONTRACT, "GBM sigma must be non-negative");
  if (d.kind == KT_MODEL_HESTON) {
    if (!(d.rho >= -1.0 && d.rho <= 1.0)) raise(KT_ERR_MALFORMED_CONTRACT, "Heston rho must be in [-1, 1]");
    if (d.v0 < 0.0) raise(KT_ERR_MALFORMED_CONTRACT, "Heston v0 must be non-negative");
    if (d.sigma_v < 0.0) raise(KT_ERR_MALFORMED_CONTRACT, "Heston sigma_v must be non-negative");
    if (d.theta < 0.0) raise(KT_ERR_MALFORMED_CONTRACT, "Heston theta must be non-negative");
    if (d.kappa < 0.0) raise(KT_ERR_MALFORMED_CONTRACT, "Heston kappa must be non-negative");
    if (d.scheme != KT_HESTON_FLOOR && d.scheme != KT_HESTON_FULL_TRUNCATION)
      raise(KT_ERR_INVALID_ARGUMENT, "unknown Heston scheme");
  }
}

int model_normals(const ModelDesc& d) {
  return d.kind == KT_MODEL_GBM ? 1 : d.kind == KT_MODEL_HESTON ? 2 : (int)d.names.size();
}

int model_kernel_kind(const ModelDesc& d) {
  return d.kind == KT_MODEL_GBM ? MK_GBM : d.kind == KT_MODEL_HESTON ? MK_HESTON : MK_BASKET;
}

// Per grid point model coefficients (see kt_launch.h).
std::vector<float> model_table(const ModelDesc& d, const std::vector<double>& grid, int* stride) {
  size_t n = grid.size();
  std::vector<float> tab;
  double prev = 0.0;
  if (d.kind == KT_MODEL_HESTON) {
    *stride = kHestonStride;
    tab.assign(n * kHestonStride, 0.0f);
    double r = d.mu[0];
    for (size_t k = 0; k < n; ++k) {
      double dt = grid[k] - prev;
      prev = grid[k];
      if (!(dt > 0.0)) continue;
      double sdt = std::sqrt(dt);
      float* t = &tab[k * kHestonStride];
      t[0] = (float)(r * dt * kLog2e);
      t[1] = (float)(-0.5 * dt * kLog2e);
      t[2] = (float)(sdt * kLog2e);
      t[3] = (float)(-d.kappa * dt);
      t[4] = (float)(d.kappa * d.theta * dt);
      t[5] = (float)(d.sigma_v * sdt);
    }
    return tab;
  }
  size_t K = d.names.size();
  *stride = (int)(2 * K);
  tab.assign(n * 2 * K, 0.0f);
  for (size_t k = 0; k < n; ++k) {
    double dt = grid[k] - prev;
    prev = grid[k];
    if (!(dt > 0.0)) continue;
    for (size_t i = 0; i < K; ++i) {
      double sg = d.sigma[i];
      tab[(k * K + i) * 2] = (float)((d.mu[i] - 0.5 * sg * sg) * dt * kLog2e);
      tab[(k * K + i) * 2 + 1] = (float)(sg * std::sqrt(dt) * kLog2e);
    }
  }
  return tab;
}

kt_price_result result_from_stats(const KStats& s) {
  kt_price_result r{};
  r.n_paths = (uint64_t)s.n;
  r.price = s.n > 0 ? s.mean : 0.0;
  r.sample_std = s.n > 1 ? std::sqrt(s.m2 / (s.n - 1.0)) : 0.0;
  r.std_error = s.n > 0 ? r.sample_std / std::sqrt(s.n) : 0.0;
  r.ci95_low = r.price - kZ95 * r.std_error;
  r.ci95_high = r.price + kZ95 * r.std_error;
  return r;
}

void check_hip(hipError_t e, const char* what) {
  if (e != hipSuccess) raise(KT_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

int lowest_error_code(unsigned bits) {
  for (int c = 1; c < 32; ++c)
    if (bits & (1u << c)) return c;
  return 0;
}

const char* error_text(int code) {
  switch (code) {
    case KT_ERR_UNKNOWN_ASSET: return "unknown asset referenced by the contract";
    case KT_ERR_TIME_OUT_OF_RANGE: return "a fixing date of AverageAt is not a grid point";
    case KT_ERR_MALFORMED_CONTRACT: return "malformed contract: empty or invalid average window or fixing set";
    case KT_ERR_UNSUPPORTED:
      return "unsupported: `or` reached on a path, or an average window extending past its evaluation step";
    default: return "error raised on a path";
  }
}

std::shared_ptr<const Program> lower_cached(kt_engine* E, const std::shared_ptr<const Con>& con,
                                            const std::vector<double>& grid, const std::vector<std::string>& names) {
  // key: the parsed contract (the parse cache keeps one per contract text;
  // the entry holds it, so its address is not reused while the entry
  // lives), the names and a 64-bit mix of the grid's bits in four
  // independent lanes (the grid itself is compared on a hit, so a collision
  // only costs a re-lowering)
  uint64_t lane[4] = {0x9E3779B97F4A7C15ull ^ grid.size(), 0x6A09E667F3BCC909ull, 0xBB67AE8584CAA73Bull,
                      0x3C6EF372FE94F82Bull};
  const size_t n4 = grid.size() & ~(size_t)3;
  auto mix = [](uint64_t h, double t) {
    uint64_t b;
    std::memcpy(&b, &t, sizeof b);
    h = (h ^ b) * 0xBF58476D1CE4E5B9ull;
    return h ^ (h >> 31);
  };
  for (size_t i = 0; i < n4; i += 4)
    for (int l = 0; l < 4; ++l) lane[l] = mix(lane[l], grid[i + l]);
  for (size_t i = n4; i < grid.size(); ++i) lane[0] = mix(lane[0], grid[i]);
  uint64_t h = lane[0];
  for (int l = 1; l < 4; ++l) {
    h = (h ^ lane[l]) * 0x94D049BB133111EBull;
    h ^= h >> 29;
  }
  const Con* cp = con.get();
  std::string key((const char*)&cp, sizeof cp);
  for (const std::string& n : names) {
    key.push_back('\0');
    key += n;
  }
  key.push_back('\0');
  key.append((const char*)&h, sizeof h);
  auto it = E->lower_cache.find(key);
  if (it != E->lower_cache.end() && it->second.grid == grid) return it->second.prog;
  auto p = std::make_shared<const Program>(lower({con.get()}, grid, names));
  if (E->lower_cache.size() >= 16384) {
    E->lower_cache.clear();
    E->adj_plan_cache.clear();
    it = E->lower_cache.end();
  }
  if (it != E->lower_cache.end()) {
    E->adj_plan_cache.erase(it->second.prog.get());
    it->second = kt_engine::Lowered{grid, p, con};
  } else {
    E->lower_cache.emplace(std::move(key), kt_engine::Lowered{grid, p, con});
  }
  return p;
}

std::shared_ptr<const AdjPlan> adjoint_plan_cached(kt_engine* E, const std::shared_ptr<const Program>& prog) {
  auto it = E->adj_plan_cache.find(prog.get());
  if (it != E->adj_plan_cache.end()) return it->second;
  std::shared_ptr<const AdjPlan> plan;
  try {
    plan = std::make_shared<const AdjPlan>(classify_adjoint(*prog));
  } catch (const Error& e) {
    if (e.code != KT_ERR_UNSUPPORTED) throw;
  }
  E->adj_plan_cache.emplace(prog.get(), plan);
  return plan;
}

std::string error_text(int code, const Program& prog) {
  std::string s = error_text(code);
  if (code == KT_ERR_TIME_OUT_OF_RANGE && !std::isnan(prog.time_payload)) {
    char buf[64];
    std::snprintf(buf, sizeof buf, ": %.17g", prog.time_payload);
    s += buf;
  }
  return s;
}

}  // namespace kt

// ---------------------------------------------------------------------------
// Engine
// ---------------------------------------------------------------------------
kt_engine::~kt_engine() {
  (void)hipSetDevice(device);
  if (pf_exec) (void)hipGraphExecDestroy(pf_exec);
  if (pf_graph) (void)hipGraphDestroy(pf_graph);
  if (lsm_exec) (void)hipGraphExecDestroy(lsm_exec);
  if (lsm_graph) (void)hipGraphDestroy(lsm_graph);
  if (stream) (void)hipStreamSynchronize(stream);
  for (DevBuf* b : {&arena, &partials, &stats, &errw, &pv, &z, &scratch, &paths, &aux, &book, &errp, &book2, &partials2,
                    &stats2, &errw2})
    if (b->p) (void)hipFree(b->p);
  for (Staging& st : staging) {
    if (st.buf.p) (void)hipHostFree(st.buf.p);
    if (st.ev) (void)hipEventDestroy(st.ev);
  }
  for (PinBuf& b : readback_buf)
    if (b.p) (void)hipHostFree(b.p);
  for (const ProfBracket& pb : prof_open) {
    (void)hipEventDestroy(pb.a);
    (void)hipEventDestroy(pb.b);
  }
  if (prof_cur) (void)hipEventDestroy(prof_cur);
  for (hipEvent_t e : prof_pool) (void)hipEventDestroy(e);
  if (stream) (void)hipStreamDestroy(stream);
}

static hipEvent_t pooled_event(std::vector<hipEvent_t>& pool) {
  if (!pool.empty()) {
    hipEvent_t e = pool.back();
    pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  check_hip(hipEventCreate(&e), "hipEventCreate");
  return e;
}

void kt_engine::prof_begin() {
  if (!profiling) return;
  if (!prof_cur) prof_cur = pooled_event(prof_pool);
  check_hip(hipEventRecord(prof_cur, stream), "hipEventRecord");
}

void kt_engine::prof_end(double path_steps) {
  if (!profiling || !prof_cur) return;
  hipEvent_t b = pooled_event(prof_pool);
  check_hip(hipEventRecord(b, stream), "hipEventRecord");
  prof_open.push_back(ProfBracket{prof_cur, b, path_steps});
  prof_cur = nullptr;
}

// Waits for every recorded bracket and adds it to the totals.
void kt_engine::prof_resolve() {
  for (const ProfBracket& pb : prof_open) {
    check_hip(hipEventSynchronize(pb.b), "hipEventSynchronize");
    float ms = 0.0f;
    check_hip(hipEventElapsedTime(&ms, pb.a, pb.b), "hipEventElapsedTime");
    prof_ms += ms;
    prof_path_steps += pb.path_steps;
    prof_launches += 1;
    prof_pool.push_back(pb.a);
    prof_pool.push_back(pb.b);
  }
  prof_open.clear();
}

void* kt_engine::ensure(DevBuf& b, size_t bytes) {
  if (bytes <= b.cap) return b.p;
  if (b.p) check_hip(hipFree(b.p), "hipFree");
  b.p = nullptr;
  b.cap = 0;
  size_t want = std::max<size_t>(bytes, 4096);
  hipError_t e = hipMalloc(&b.p, want);
  if (e == hipErrorOutOfMemory) {
    (void)hipGetLastError();
    b.p = nullptr;
    raise(KT_ERR_OUT_OF_MEMORY, "device allocation of " + std::to_string(want) + " bytes failed");
  }
  check_hip(e, "hipMalloc");
  b.cap = want;
  return b.p;
}

static char* grow_pinned(PinBuf& b, size_t bytes) {
  if (bytes <= b.cap) return b.p;
  if (b.p) check_hip(hipHostFree(b.p), "hipHostFree");
  b.p = nullptr;
  b.cap = 0;
  const size_t want = std::max<size_t>(bytes, (size_t)1 << 16);
  hipError_t e = hipHostMalloc((void**)&b.p, want, hipHostMallocDefault);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    b.p = nullptr;
    raise(KT_ERR_OUT_OF_MEMORY, "page-locked host allocation of " + std::to_string(want) + " bytes failed");
  }
  b.cap = want;
  return b.p;
}

char* kt_engine::stage(size_t bytes, int slot) {
  Staging& st = staging[slot];
  if (st.busy) {
    check_hip(hipEventSynchronize(st.ev), "hipEventSynchronize");
    st.busy = false;
  }
  return grow_pinned(st.buf, std::max<size_t>(bytes, 256));
}

void kt_engine::stage_copy(void* dev, size_t bytes, int slot, size_t offset) {
  Staging& st = staging[slot];
  if (!st.ev) check_hip(hipEventCreateWithFlags(&st.ev, hipEventDisableTiming), "hipEventCreate");
  check_hip(hipMemcpyAsync((char*)dev + offset, st.buf.p + offset, bytes, hipMemcpyHostToDevice, stream), "upload");
  check_hip(hipEventRecord(st.ev, stream), "hipEventRecord");
  st.busy = true;
}

char* kt_engine::readback(size_t bytes, int slot) {
  return grow_pinned(readback_buf[slot], std::max<size_t>(bytes, 256));
}

// Upload a set of host blobs into the arena; returns device pointers.
std::vector<void*> kt_engine::upload(const std::vector<std::pair<const void*, size_t>>& blobs) {
  size_t total = 0;
  std::vector<size_t> offs;
  for (auto& b : blobs) {
    offs.push_back(total);
    total += (b.second + 255) & ~(size_t)255;
  }
  total = std::max<size_t>(total, 256);
  char* host = stage(total);
  for (size_t i = 0; i < blobs.size(); ++i)
    if (blobs[i].second) memcpy(host + offs[i], blobs[i].first, blobs[i].second);
  char* dev = (char*)ensure(arena, total);
  stage_copy(dev, total);
  std::vector<void*> out;
  for (size_t o : offs) out.push_back(dev + o);
  return out;
}

namespace kt {

// Run one program over n_paths paths under one or kScenarios scenarios (same
// model kind, different parameters and discount rates); returns the
// per-output statistics.  pv_host / z_host (nullable) request the debug
// variant (one scenario only).
void fill_model_args(PathArgs& A, const ModelDesc& md, uint64_t seed) {
  const int K = (int)md.names.size();
  A.heston_full_truncation = md.scheme == KT_HESTON_FULL_TRUNCATION;
  A.seed_lo = (uint32_t)seed;
  A.seed_hi = (uint32_t)(seed >> 32);
  for (int i = 0; i < K && i < kMaxAssets; ++i) A.s0[i] = (float)md.s0[i];
  for (size_t i = 0; i < md.chol.size() && i < (size_t)kMaxAssets * kMaxAssets; ++i) A.chol[i] = (float)md.chol[i];
  A.v0 = (float)std::max(md.v0, 0.0);
  A.rho = (float)md.rho;
  A.rho_perp = (float)std::sqrt(std::max(1.0 - md.rho * md.rho, 0.0));
  A.s0_scen[0] = (float)md.s0[0];
}

// exp(-rate t_k) at every grid point whose op list pays a flow (OP_EMIT, the
// only op that reads the discount table) and NaN elsewhere, so that a read
// of an unused entry would show; out has grid.size() entries.
void discount_factors(const Program& prog, const std::vector<double>& grid, double rate, double* out) {
  const double nan = std::numeric_limits<double>::quiet_NaN();
  for (size_t k = 0; k < grid.size(); ++k) {
    out[k] = nan;
    if (k >= prog.steps.size()) continue;
    const KStep& st = prog.steps[k];
    for (int i = 0; i < st.count; ++i)
      if (prog.ops[(size_t)st.begin + i].code == OP_EMIT) {
        out[k] = std::exp(-rate * grid[k]);
        break;
      }
  }
}

// Knock-out thresholds log2(H/s0) per group and scenario; groups whose spot
// or level is not positive are interpreted instead.
std::vector<KGroup> knockout_tables(const Program& prog, const std::vector<ModelDesc>& models,
                                    std::vector<float>* thr) {
  const size_t SC = models.size();
  const ModelDesc& md = models[0];
  std::vector<KGroup> groups = prog.groups;
  thr->assign(std::max<size_t>(groups.size() * SC, 1), 0.0f);
  for (size_t g = 0; g < groups.size(); ++g) {
    KGroup& gd = groups[g];
    if (gd.kind < G_KO_GE || gd.kind > G_KO_LT) continue;
    bool ok = gd.imm > 0.0f && gd.asset >= 0 && gd.asset < (int)md.s0.size();
    for (size_t sc = 0; ok && sc < SC; ++sc) {
      double s0 = models[sc].s0[gd.asset];
      ok = s0 > 0.0;
      if (ok) (*thr)[g * SC + sc] = (float)std::log2((double)gd.imm / s0);
    }
    if (!ok) gd.kind = G_GENERIC;
  }
  return groups;
}

LaunchTables launch_tables(const Program& prog, const std::vector<ModelDesc>& models,
                           const std::vector<double>& rates, const std::vector<double>& grid, bool discount) {
  const int SC = (int)models.size();
  const ModelDesc& md = models[0];
  LaunchTables T;
  if (SC == 1) {
    T.tab = model_table(md, grid, &T.stride);
  } else {
    if (md.kind != KT_MODEL_GBM) raise(KT_ERR_UNSUPPORTED, "scenario kernel is GBM only");
    // A GBM table depends on (mu, sigma) only: spot-bumped scenarios share it.
    std::vector<std::vector<float>> per;
    for (int sc = 0; sc < SC; ++sc) {
      int same = -1;
      for (int q = 0; q < sc && same < 0; ++q)
        if (models[q].mu == models[sc].mu && models[q].sigma == models[sc].sigma) same = q;
      per.push_back(same >= 0 ? per[same] : model_table(models[sc], grid, &T.stride));
    }
    T.tab.resize(grid.size() * SC * 2);
    for (size_t k = 0; k < grid.size(); ++k)
      for (int sc = 0; sc < SC; ++sc) {
        T.tab[(k * SC + sc) * 2] = per[sc][k * 2];
        T.tab[(k * SC + sc) * 2 + 1] = per[sc][k * 2 + 1];
      }
    T.stride = 2 * SC;
  }
  // discount tables, one per distinct rate
  std::vector<double> distinct;
  T.df_index.assign(SC, 0);
  for (int sc = 0; sc < SC; ++sc) {
    auto it = std::find(distinct.begin(), distinct.end(), rates[sc]);
    if (it == distinct.end()) {
      T.df_index[sc] = (int)distinct.size();
      distinct.push_back(rates[sc]);
    } else {
      T.df_index[sc] = (int)(it - distinct.begin());
    }
  }
  if (discount) {
    T.df.resize(distinct.size() * grid.size());
    for (size_t t = 0; t < distinct.size(); ++t) discount_factors(prog, grid, distinct[t], &T.df[t * grid.size()]);
  }
  T.groups = knockout_tables(prog, models, &T.thr);
  return T;
}

std::vector<std::pair<const void*, size_t>> program_blobs(const Program& prog, const LaunchTables& T,
                                                         std::vector<KOp>* ops) {
  *ops = prog.ops;
  if (ops->empty()) ops->push_back(KOp{OP_NOP, 0, 0, 0, 0.0f});
  return {{ops->data(), ops->size() * sizeof(KOp)},
          {prog.steps.data(), prog.steps.size() * sizeof(KStep)},
          {T.groups.data(), T.groups.size() * sizeof(KGroup)},
          {prog.init.data(), prog.init.size() * sizeof(KOp)},
          {T.df.data(), T.df.size() * sizeof(double)},
          {T.tab.data(), T.tab.size() * sizeof(float)},
          {T.thr.data(), T.thr.size() * sizeof(float)}};
}

PathArgs program_path_args(const Program& prog, const std::vector<ModelDesc>& models, const LaunchTables& T,
                           const std::vector<double>& grid, uint64_t seed, const std::vector<void*>& dev) {
  const int SC = (int)models.size();
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
  A.n_groups = (int)T.groups.size();
  A.n_uniform = uniform_group_prefix(T.groups.data(), A.n_groups);
  A.serial_tail = scenario_serial_tail(T.groups.data(), A.n_groups);
  A.tab_uniform = uniform_table_rows(T.tab, grid.size(), (size_t)T.stride);
  A.tab_stride = T.stride;
  fill_model_args(A, models[0], seed);
  for (int sc = 0; sc < SC && sc < kScenarios; ++sc) {
    A.s0_scen[sc] = (float)models[sc].s0[0];
    A.df_index[sc] = T.df_index[sc];
  }
  return A;
}

std::vector<KStats> run_program_scen(kt_engine* E, const Program& prog, const std::vector<ModelDesc>& models,
                                     const std::vector<double>& rates, const std::vector<std::vector<double>>& stat_w,
                                     const std::vector<double>& grid, uint64_t n_paths, uint64_t seed,
                                     double* pv_host, float* z_host, double* pv_keep, bool anti, int combine) {
  const int SC = (int)models.size();
  const ModelDesc& md = models[0];
  LaunchTables T = launch_tables(prog, models, rates, grid);

  int kind = model_kernel_kind(md);
  int K = (int)md.names.size();
  bool debug = pv_host || z_host || pv_keep;
  int ns = 0, nout = 0;
  if (!path_variant_exists(kind, K, prog.n_slots, prog.n_outputs, debug, SC, &ns, &nout, anti))
    raise(KT_ERR_UNSUPPORTED, "no path kernel variant for this program shape");
  if (combine < 0 || combine > std::min(prog.n_outputs, kScenarios) || (combine > 0 && (SC != 1 || nout < 2)) ||
      (combine > 0 && (stat_w.empty() || (int)stat_w[0].size() < combine)))
    raise(KT_ERR_INVALID_ARGUMENT, "bad output combination");
  if (prog.n_slots > ns) raise(KT_ERR_UNSUPPORTED, "program needs more registers than the kernel variant holds");
  const int nstat = SC == 1 ? nout : kScenarioStats;
  const int n_used = SC == 1 ? prog.n_outputs : kScenarioStats;

  std::vector<KOp> ops;
  auto dev = E->upload(program_blobs(prog, T, &ops));

  const uint64_t kMaxBlocksPerLaunch = 1ull << 20;
  uint64_t total_blocks = (n_paths + kBlock - 1) / kBlock;
  if (total_blocks > (1ull << 31)) raise(KT_ERR_INVALID_ARGUMENT, "too many paths");
  KStats* partials = (KStats*)E->ensure(E->partials, std::max<uint64_t>(total_blocks, 1) * nstat * sizeof(KStats));
  // [n_used statistics][error word copy]: one readback
  KStats* stats = (KStats*)E->ensure(E->stats, (kScenarioStats + 8 + 1) * sizeof(KStats));
  unsigned* err_copy = (unsigned*)(stats + n_used);
  unsigned* err = (unsigned*)E->ensure(E->errp, 256);
  if (!E->errp_zero) check_hip(hipMemsetAsync(err, 0, sizeof(unsigned), E->stream), "memset");
  E->errp_zero = false;  // until the finalize below has cleared it again
  double* pv_dev = nullptr;
  float* z_dev = nullptr;
  int M = model_normals(md);
  if (pv_keep) pv_dev = pv_keep;
  else if (pv_host) pv_dev = (double*)E->ensure(E->pv, std::max<uint64_t>(n_paths, 1) * sizeof(double));
  if (z_host) z_dev = (float*)E->ensure(E->z, std::max<uint64_t>(n_paths * grid.size() * M, 1) * sizeof(float));

  PathArgs A = program_path_args(prog, models, T, grid, seed, dev);
  if (combine > 0) {
    for (int o = 0; o < combine; ++o) A.stat_w[0][o] = stat_w[0][o];
    A.combine = combine;
  } else {
    for (size_t j = 0; j < stat_w.size() && j < (size_t)kScenarioStats; ++j)
      for (int sc = 0; sc < SC && sc < kScenarios; ++sc) A.stat_w[j][sc] = stat_w[j][sc];
  }
  A.err = err;

  E->prof_begin();
  for (uint64_t b0 = 0; b0 < total_blocks; b0 += kMaxBlocksPerLaunch) {
    uint64_t nb = std::min(kMaxBlocksPerLaunch, total_blocks - b0);
    A.path_base = b0 * kBlock;
    A.n_paths = std::min<uint64_t>(nb * kBlock, n_paths - b0 * kBlock);
    A.partials = partials + b0 * nstat;
    A.pv_out = pv_dev ? pv_dev + b0 * kBlock : nullptr;
    A.z_out = z_dev ? z_dev + b0 * kBlock * grid.size() * M : nullptr;
    check_hip(launch_path(kind, K, ns, nout, debug, SC, A, (int)nb, E->stream, anti), "path kernel launch");
  }
  E->prof_end((double)n_paths * (double)(grid.size() > 0 ? grid.size() - 1 : 0) * SC * (anti ? 2 : 1));
  check_hip(launch_finalize(partials, (int)total_blocks, nstat, n_used, stats, E->stream, err, err_copy),
            "finalize launch");
  const size_t rb_bytes = (n_used + 1) * sizeof(KStats);
  char* rb = E->readback(rb_bytes);
  check_hip(hipMemcpyAsync(rb, stats, rb_bytes, hipMemcpyDeviceToHost, E->stream), "copy stats");
  if (pv_host && pv_dev)
    check_hip(hipMemcpyAsync(pv_host, pv_dev, n_paths * sizeof(double), hipMemcpyDeviceToHost, E->stream), "copy pv");
  if (z_host)
    check_hip(hipMemcpyAsync(z_host, z_dev, n_paths * grid.size() * M * sizeof(float), hipMemcpyDeviceToHost,
                             E->stream),
              "copy normals");
  check_hip(hipStreamSynchronize(E->stream), "stream sync");
  E->errp_zero = true;
  std::vector<KStats> out(n_used);
  memcpy(out.data(), rb, n_used * sizeof(KStats));
  unsigned err_bits = 0;
  memcpy(&err_bits, rb + n_used * sizeof(KStats), sizeof(unsigned));
  if (err_bits) {
    int code = lowest_error_code(err_bits);
    raise(code, error_text(code, prog));
  }
  return out;
}

std::vector<KStats> run_program(kt_engine* E, const Program& prog, const ModelDesc& md,
                                const std::vector<double>& grid, double rate, uint64_t n_paths, uint64_t seed,
                                double* pv_host, float* z_host, double* pv_keep) {
  return run_program_scen(E, prog, {md}, {rate}, {}, grid, n_paths, seed, pv_host, z_host, pv_keep);
}

// Shared front half of price_gbm: parse, plan, grid, validation.
PreparedContract prepare(const char* json, const ModelDesc& md, const kt_mc_config* cfg, bool allow_vr,
                         kt_engine* E) {
  if (!cfg) raise(KT_ERR_INVALID_ARGUMENT, "null config");
  if (!allow_vr && (cfg->antithetic || cfg->control_variate))
    raise(KT_ERR_UNSUPPORTED, "variance reduction is not available for this entry point");
  PreparedContract pc;
  Plan plan_local;
  const Plan* plan_p = &plan_local;
  if (E) {
    // a text that fails to parse raises here and is never cached
    auto it = E->parse_cache.find(json);
    if (it == E->parse_cache.end()) {
      kt_engine::Parsed entry;
      entry.con = parse_contract(json);
      entry.plan = compile_plan(*entry.con);
      if (E->parse_cache.size() >= 16384) E->parse_cache.clear();
      it = E->parse_cache.emplace(std::string(json), std::move(entry)).first;
    }
    pc.con = it->second.con;
    plan_p = &it->second.plan;
  } else {
    pc.con = parse_contract(json);
    plan_local = compile_plan(*pc.con);
  }
  const Plan& plan = *plan_p;
  if (cfg->grid && cfg->n_grid) {
    pc.grid.assign(cfg->grid, cfg->grid + cfg->n_grid);
  } else if (cfg->grid == nullptr && cfg->n_grid == 0) {
    pc.grid = time_grid_forced(plan, cfg->steps_per_year, md.kind == KT_MODEL_HESTON);
  } else {
    raise(KT_ERR_INCONSISTENT_PATH, "empty time grid");
  }
  validate_model(md);
  validate_grid(pc.grid);
  return pc;
}

}  // namespace kt

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
namespace kt {
int set_error(kt_engine* E, int code, const std::string& msg) {
  if (E) E->err = msg;
  g_last_error = msg;
  return code;
}
}  // namespace kt

namespace {
// Average time of `reps` launches after one warm launch, on the engine's stream.
struct ProbeEvents {
  hipEvent_t a = nullptr, b = nullptr;
  ~ProbeEvents() {
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
  }
};

template <class F>
double time_probe(kt_engine* E, int reps, F launch) {
  check_hip(launch(), "probe launch");
  ProbeEvents ev;  // destroyed on every exit, the error paths included
  check_hip(hipEventCreate(&ev.a), "event");
  check_hip(hipEventCreate(&ev.b), "event");
  check_hip(hipEventRecord(ev.a, E->stream), "record");
  for (int i = 0; i < reps; ++i) check_hip(launch(), "probe launch");
  check_hip(hipEventRecord(ev.b, E->stream), "record");
  check_hip(hipEventSynchronize(ev.b), "sync");
  float t = 0.0f;
  check_hip(hipEventElapsedTime(&t, ev.a, ev.b), "elapsed");
  return (double)t / reps;
}
}  // namespace

extern "C" {

kt_status kt_engine_create(int device, kt_engine** out) {
  if (!out) return (kt_status)set_error(nullptr, KT_ERR_INVALID_ARGUMENT, "null out");
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
    return (kt_status)set_error(nullptr, KT_ERR_DEVICE, "no HIP device available");
  if (device < 0 || device >= n) return (kt_status)set_error(nullptr, KT_ERR_INVALID_ARGUMENT, "bad device ordinal");
  if (hipSetDevice(device) != hipSuccess) return (kt_status)set_error(nullptr, KT_ERR_DEVICE, "hipSetDevice failed");
  kt_engine* E = new kt_engine();
  E->device = device;
  // the allocation-free entry point's scratch address on this device (its
  // failure is reported by that entry point, not here)
  (void)kt::noalloc_scratch_init(device);
  if (hipStreamCreateWithFlags(&E->stream, hipStreamNonBlocking) != hipSuccess) {
    delete E;
    return (kt_status)set_error(nullptr, KT_ERR_DEVICE, "hipStreamCreate failed");
  }
  *out = E;
  return KT_OK;
}

void kt_engine_destroy(kt_engine* engine) { delete engine; }

const char* kt_last_error(const kt_engine* engine) { return engine ? engine->err.c_str() : ""; }
const char* kt_last_error_global(void) { return g_last_error.c_str(); }

kt_status kt_price(kt_engine* E, const char* json, const kt_model* model, const kt_mc_config* cfg,
                   kt_price_result* out) {
  KT_GUARD_BEGIN
  if (!E || !out) raise(KT_ERR_INVALID_ARGUMENT, "null engine or output");
  check_hip(hipSetDevice(E->device), "hipSetDevice");
  ModelDesc md = model_from_abi(model);
  PreparedContract pc = prepare(json, md, cfg, true);
  *out = kt_price_result{};
  if (cfg->antithetic || cfg->control_variate) {
    *out = price_vr(E, *pc.con, md, pc.grid, cfg->rate, cfg->n_paths, cfg->seed, cfg->antithetic != 0,
                    cfg->control_variate != 0);
    E->err.clear();
    return KT_OK;
  }
  if (cfg->n_paths == 0) return KT_OK;
  Program prog = lower({pc.con.get()}, pc.grid, md.names);
  auto st = run_program(E, prog, md, pc.grid, cfg->rate, cfg->n_paths, cfg->seed, nullptr, nullptr);
  *out = result_from_stats(st[0]);
  E->err.clear();
  return KT_OK;
  KT_GUARD_END(E)
}

kt_status kt_debug_paths(kt_engine* E, const char* json, const kt_model* model, const kt_mc_config* cfg,
                         double* pv_out, float* normals_out, double* grid_out, uint64_t grid_cap, uint64_t* n_grid,
                         kt_price_result* out) {
  KT_GUARD_BEGIN
  if (!E || !out || !pv_out) raise(KT_ERR_INVALID_ARGUMENT, "null engine or output");
  check_hip(hipSetDevice(E->device), "hipSetDevice");
  ModelDesc md = model_from_abi(model);
  PreparedContract pc = prepare(json, md, cfg);
  if (n_grid) *n_grid = pc.grid.size();
  if (grid_out)
    for (size_t i = 0; i < pc.grid.size() && i < grid_cap; ++i) grid_out[i] = pc.grid[i];
  *out = kt_price_result{};
  if (cfg->n_paths == 0) return KT_OK;
  Program prog = lower({pc.con.get()}, pc.grid, md.names);
  auto st = run_program(E, prog, md, pc.grid, cfg->rate, cfg->n_paths, cfg->seed, pv_out, normals_out);
  *out = result_from_stats(st[0]);
  return KT_OK;
  KT_GUARD_END(E)
}

kt_status kt_price_batch(kt_engine* E, const char* const* jsons, uint64_t n_contracts, const kt_model* model,
                         const kt_mc_config* cfg, kt_price_result* out) {
  KT_GUARD_BEGIN
  if (!E) raise(KT_ERR_INVALID_ARGUMENT, "null engine");
  if (n_contracts == 0) return KT_OK;
  if (!jsons || !out || !cfg) raise(KT_ERR_INVALID_ARGUMENT, "null argument");
  // Like price_batch_gbm, the batch simulates plain paths whatever the
  // variance-reduction flags say (src/pricer.rs:250-294 never reads them).
  check_hip(hipSetDevice(E->device), "hipSetDevice");
  ModelDesc md = model_from_abi(model);
  // Merged plan (src/pricer.rs:260-278).
  std::vector<std::unique_ptr<Con>> cons;
  Plan merged;
  for (uint64_t i = 0; i < n_contracts; ++i) {
    cons.push_back(parse_contract(jsons[i]));
    Plan p = compile_plan(*cons.back());
    merged.fixed_dates.insert(merged.fixed_dates.end(), p.fixed_dates.begin(), p.fixed_dates.end());
    merged.horizon = std::max(merged.horizon, p.horizon);
    merged.needs_fine_grid = merged.needs_fine_grid || p.needs_fine_grid;
  }
  std::vector<double> grid = time_grid_forced(merged, cfg->steps_per_year, md.kind == KT_MODEL_HESTON);
  validate_model(md);
  validate_grid(grid);
  if (cfg->n_paths == 0) {
    for (uint64_t i = 0; i < n_contracts; ++i) out[i] = kt_price_result{};
    return KT_OK;
  }
  // Up to kMaxOutputs contracts share one launch; a group that does not fit
  // the register file is split.
  std::vector<std::pair<size_t, size_t>> todo;
  for (size_t b = 0; b < n_contracts; b += kMaxOutputs) todo.push_back({b, std::min<size_t>(b + kMaxOutputs, n_contracts)});
  while (!todo.empty()) {
    auto [b, e] = todo.back();
    todo.pop_back();
    std::vector<const Con*> group;
    for (size_t i = b; i < e; ++i) group.push_back(cons[i].get());
    Program prog;
    try {
      prog = lower(group, grid, md.names);
    } catch (const Error& err) {
      if (err.code == KT_ERR_UNSUPPORTED && e - b > 1 && err.msg.find("registers") != std::string::npos) {
        size_t mid = b + (e - b) / 2;
        todo.push_back({mid, e});
        todo.push_back({b, mid});
        continue;
      }
      throw;
    }
    auto st = run_program(E, prog, md, grid, cfg->rate, cfg->n_paths, cfg->seed, nullptr, nullptr);
    for (size_t i = b; i < e; ++i) out[i] = result_from_stats(st[i - b]);
  }
  E->err.clear();
  return KT_OK;
  KT_GUARD_END(E)
}

kt_status kt_greeks_gbm(kt_engine* E, const char* json, const char* asset, double s0, double sigma,
                        const kt_mc_config* cfg, const kt_bumps* bumps, kt_greeks* out) {
  KT_GUARD_BEGIN
  if (!E || !out || !cfg || !asset) raise(KT_ERR_INVALID_ARGUMENT, "null argument");
  check_hip(hipSetDevice(E->device), "hipSetDevice");
  kt_bumps h = bumps ? *bumps : kt_bumps{1e-2, 1e-2, 1e-4};
  greeks_gbm_crn(E, json, asset, s0, sigma, cfg, h, out);
  E->err.clear();
  return KT_OK;
  KT_GUARD_END(E)
}

kt_status kt_greek_surface(kt_engine* E, const char* json, const char* asset, const double* spots, uint64_t n_spots,
                           const double* vols, uint64_t n_vols, const kt_mc_config* cfg, const kt_bumps* bumps,
                           kt_greeks* out) {
  KT_GUARD_BEGIN
  if (!E || !cfg || !asset || !json) raise(KT_ERR_INVALID_ARGUMENT, "null argument");
  if (!out && n_spots && n_vols) raise(KT_ERR_INVALID_ARGUMENT, "null output");
  check_hip(hipSetDevice(E->device), "hipSetDevice");
  kt_bumps h = bumps ? *bumps : kt_bumps{1e-2, 1e-2, 1e-4};
  greek_surface(E, json, asset, spots, n_spots, vols, n_vols, cfg, h, out);
  E->err.clear();
  return KT_OK;
  KT_GUARD_END(E)
}

kt_status kt_greeks_bump(kt_engine* E, const char* json, const kt_model* model, const kt_mc_config* cfg,
                         const double* h, uint64_t n_params, double* out, double* stderr_out) {
  KT_GUARD_BEGIN
  if (!E || !out || !cfg || !h) raise(KT_ERR_INVALID_ARGUMENT, "null argument");
  check_hip(hipSetDevice(E->device), "hipSetDevice");
  greeks_bump(E, json, model, cfg, h, n_params, out, stderr_out);
  E->err.clear();
  return KT_OK;
  KT_GUARD_END(E)
}

kt_status kt_greeks_adjoint(kt_engine* E, const char* json, const kt_model* model, const kt_mc_config* cfg,
                            double* out, double* stderr_out, uint64_t n_out) {
  KT_GUARD_BEGIN
  if (!E || !out || !cfg) raise(KT_ERR_INVALID_ARGUMENT, "null argument");
  check_hip(hipSetDevice(E->device), "hipSetDevice");
  greeks_adjoint(E, json, model, cfg, out, stderr_out, n_out);
  E->err.clear();
  return KT_OK;
  KT_GUARD_END(E)
}

kt_status kt_price_american(kt_engine* E, const char* json, const double* dates, uint64_t n_dates,
                            const kt_model* model, const kt_mc_config* cfg, int32_t n_basis, kt_price_result* out) {
  KT_GUARD_BEGIN
  if (!E || !out || !cfg) raise(KT_ERR_INVALID_ARGUMENT, "null argument");
  check_hip(hipSetDevice(E->device), "hipSetDevice");
  price_american(E, json, dates, n_dates, model, cfg, n_basis, out);
  E->err.clear();
  return KT_OK;
  KT_GUARD_END(E)
}

kt_status kt_debug_american(kt_engine* E, const char* json, const double* dates, uint64_t n_dates,
                            const kt_model* model, const kt_mc_config* cfg, int32_t n_basis, double* pv_out,
                            double* coef_out, kt_price_result* out) {
  KT_GUARD_BEGIN
  if (!E || !out || !cfg || !pv_out) raise(KT_ERR_INVALID_ARGUMENT, "null argument");
  check_hip(hipSetDevice(E->device), "hipSetDevice");
  price_american(E, json, dates, n_dates, model, cfg, n_basis, out, pv_out, coef_out);
  E->err.clear();
  return KT_OK;
  KT_GUARD_END(E)
}

kt_status kt_engine_set_option(kt_engine* E, const char* name, int64_t value) {
  KT_GUARD_BEGIN
  if (!E || !name) raise(KT_ERR_INVALID_ARGUMENT, "null argument");
  const std::string key(name);
  if (key == "adjoint_segment") {
    if (value != 16 && value != 32) raise(KT_ERR_INVALID_ARGUMENT, "adjoint_segment must be 16 or 32");
    E->adj_seg = (int)value;
  } else if (key == "portfolio_blocks_per_cu") {
    if (value < 0 || value > 64) raise(KT_ERR_INVALID_ARGUMENT, "portfolio_blocks_per_cu must be in [0, 64]");
    E->pf_blocks_per_cu = (int)value;
  } else if (key == "portfolio_chunk") {
    if (value < 1 || value > 1024) raise(KT_ERR_INVALID_ARGUMENT, "portfolio_chunk must be in [1, 1024]");
    E->pf_chunk = (int)value;
  } else if (key == "pathwise_mode") {
    if (value != 0 && value != 1) raise(KT_ERR_INVALID_ARGUMENT, "pathwise_mode must be 0 (reverse) or 1 (tangent)");
    E->pathwise_tangent = value == 1;
  } else if (key == "portfolio_graph") {
    if (value != 0 && value != 1) raise(KT_ERR_INVALID_ARGUMENT, "portfolio_graph must be 0 or 1");
    E->pf_use_graph = value == 1;
  } else if (key == "portfolio_waves") {
    if (value != 0 && value != 6 && value != 8) raise(KT_ERR_INVALID_ARGUMENT, "portfolio_waves must be 0, 6 or 8");
    E->pf_tile.waves = (int)value;
  } else if (key == "portfolio_unroll") {
    if (value != 1 && value != 2) raise(KT_ERR_INVALID_ARGUMENT, "portfolio_unroll must be 1 or 2");
    E->pf_tile.unroll = (int)value;
  } else if (key == "portfolio_lanes") {
    if (value != 128 && value != 256) raise(KT_ERR_INVALID_ARGUMENT, "portfolio_lanes must be 128 or 256");
    E->pf_tile.lanes = (int)value;
  } else if (key == "portfolio_adjoint_waves") {
    if (value != 0 && value != 5) raise(KT_ERR_INVALID_ARGUMENT, "portfolio_adjoint_waves must be 0 or 5");
    E->pf_adj_waves = (int)value;
  } else if (key == "lsm_graph") {
    if (value != 0 && value != 1) raise(KT_ERR_INVALID_ARGUMENT, "lsm_graph must be 0 or 1");
    E->lsm_use_graph = value == 1;
  } else {
    raise(KT_ERR_INVALID_ARGUMENT, "unknown engine option: " + key);
  }
  E->err.clear();
  return KT_OK;
  KT_GUARD_END(E)
}

kt_status kt_price_portfolio(kt_engine* E, const kt_job* jobs, uint64_t n_jobs, kt_price_result* out) {
  KT_GUARD_BEGIN
  if (!E) raise(KT_ERR_INVALID_ARGUMENT, "null engine");
  if (n_jobs == 0) return KT_OK;
  if (!jobs || !out) raise(KT_ERR_INVALID_ARGUMENT, "null argument");
  check_hip(hipSetDevice(E->device), "hipSetDevice");
  price_portfolio(E, jobs, n_jobs, out);
  E->err.clear();
  return KT_OK;
  KT_GUARD_END(E)
}

kt_status kt_price_portfolio_greeks(kt_engine* E, const kt_job* jobs, uint64_t n_jobs, kt_price_result* out,
                                    double* greeks, double* greeks_se, uint64_t stride) {
  KT_GUARD_BEGIN
  if (!E) raise(KT_ERR_INVALID_ARGUMENT, "null engine");
  if (n_jobs == 0) return KT_OK;
  check_hip(hipSetDevice(E->device), "hipSetDevice");
  price_portfolio_greeks(E, jobs, n_jobs, out, greeks, greeks_se, stride);
  E->err.clear();
  return KT_OK;
  KT_GUARD_END(E)
}

kt_status kt_debug_philox(kt_engine* E, const uint32_t* ctr, const uint32_t* key, uint64_t n, uint32_t* out) {
  KT_GUARD_BEGIN
  if (!E || !ctr || !key || !out) raise(KT_ERR_INVALID_ARGUMENT, "null argument");
  check_hip(hipSetDevice(E->device), "hipSetDevice");
  if (n == 0) return KT_OK;
  uint32_t* d = (uint32_t*)E->ensure(E->scratch, n * 32);
  check_hip(hipMemcpyAsync(d, ctr, n * 16, hipMemcpyHostToDevice, E->stream), "copy");
  check_hip(launch_debug_philox(d, key[0], key[1], n, d + 4 * n, E->stream), "launch");
  check_hip(hipMemcpyAsync(out, d + 4 * n, n * 16, hipMemcpyDeviceToHost, E->stream), "copy");
  check_hip(hipStreamSynchronize(E->stream), "sync");
  return KT_OK;
  KT_GUARD_END(E)
}

kt_status kt_debug_copy(kt_engine* E, uint64_t bytes, int32_t reps, double* ms) {
  KT_GUARD_BEGIN
  if (!E || bytes % 16 || reps < 1) raise(KT_ERR_INVALID_ARGUMENT, "bad copy arguments");
  check_hip(hipSetDevice(E->device), "hipSetDevice");
  char* buf = (char*)E->ensure(E->aux, 2 * bytes);
  check_hip(hipMemsetAsync(buf, 1, bytes, E->stream), "memset");
  const double avg_ms = time_probe(E, reps, [&] { return launch_debug_copy(buf, buf + bytes, bytes, E->stream); });
  if (ms) *ms = avg_ms;
  return KT_OK;
  KT_GUARD_END(E)
}

kt_status kt_debug_valu(kt_engine* E, int32_t iters, int32_t reps, double* ms, double* lane_ops) {
  KT_GUARD_BEGIN
  if (!E || iters < 8 || iters % 8 || reps < 1) raise(KT_ERR_INVALID_ARGUMENT, "bad VALU probe arguments");
  check_hip(hipSetDevice(E->device), "hipSetDevice");
  float* out = (float*)E->ensure(E->aux, (size_t)kValuProbeBlocks * 256 * sizeof(float));
  const double avg_ms = time_probe(E, reps, [&] { return launch_debug_valu(out, iters, E->stream); });
  if (ms) *ms = avg_ms;
  if (lane_ops) *lane_ops = (double)kValuProbeBlocks * 256 * (double)iters * 8.0 / (avg_ms * 1e-3);
  return KT_OK;
  KT_GUARD_END(E)
}

kt_status kt_debug_rng_rate(kt_engine* E, int32_t normals, int32_t iters, int32_t reps, double* ms,
                            double* words) {
  KT_GUARD_BEGIN
  if (!E || iters < 1 || reps < 1) raise(KT_ERR_INVALID_ARGUMENT, "bad RNG probe arguments");
  check_hip(hipSetDevice(E->device), "hipSetDevice");
  float* out = (float*)E->ensure(E->aux, (size_t)kValuProbeBlocks * 256 * sizeof(float));
  const double avg_ms =
      time_probe(E, reps, [&] { return launch_debug_rng_rate(normals != 0, out, iters, E->stream); });
  if (ms) *ms = avg_ms;
  if (words) *words = (double)kValuProbeBlocks * 256 * (double)iters * 4.0 / (avg_ms * 1e-3);
  return KT_OK;
  KT_GUARD_END(E)
}

kt_status kt_debug_normals(kt_engine* E, const uint32_t* words, uint64_t n, float* out) {
  KT_GUARD_BEGIN
  if (!E || !words || !out) raise(KT_ERR_INVALID_ARGUMENT, "null argument");
  check_hip(hipSetDevice(E->device), "hipSetDevice");
  if (n == 0) return KT_OK;
  uint32_t* d = (uint32_t*)E->ensure(E->scratch, n * 8);
  check_hip(hipMemcpyAsync(d, words, n * 4, hipMemcpyHostToDevice, E->stream), "copy");
  check_hip(launch_debug_normals(d, n, (float*)(d + n), E->stream), "launch");
  check_hip(hipMemcpyAsync(out, d + n, n * 4, hipMemcpyDeviceToHost, E->stream), "copy");
  check_hip(hipStreamSynchronize(E->stream), "sync");
  return KT_OK;
  KT_GUARD_END(E)
}

kt_status kt_engine_set_profiling(kt_engine* E, int32_t enabled) {
  if (!E) return (kt_status)set_error(nullptr, KT_ERR_INVALID_ARGUMENT, "null engine");
  E->profiling = enabled != 0;
  return KT_OK;
}

kt_status kt_engine_kernel_time(kt_engine* E, double* total_ms, uint64_t* n_launches, double* path_steps,
                                int32_t reset) {
  if (!E) return (kt_status)set_error(nullptr, KT_ERR_INVALID_ARGUMENT, "null engine");
  KT_GUARD_BEGIN
  check_hip(hipSetDevice(E->device), "hipSetDevice");
  E->prof_resolve();
  KT_GUARD_END(E)
  if (total_ms) *total_ms = E->prof_ms;
  if (n_launches) *n_launches = E->prof_launches;
  if (path_steps) *path_steps = E->prof_path_steps;
  if (reset) {
    E->prof_ms = 0.0;
    E->prof_launches = 0;
    E->prof_path_steps = 0.0;
  }
  return KT_OK;
}

double kt_black_scholes_call(double s, double k, double t, double r, double sigma) {
  return bs_call_reference(s, k, t, r, sigma);
}

kt_status kt_plan(const char* json, uint64_t spy, int32_t force_fine, double* grid, uint64_t* n_grid, double* dates,
                  uint64_t* n_dates, uint64_t cap, int32_t* needs_fine, double* horizon, char* assets,
                  uint64_t assets_cap) {
  KT_GUARD_BEGIN
  auto c = parse_contract(json);
  Plan p = compile_plan(*c);
  auto g = time_grid_forced(p, spy, force_fine != 0);
  if (n_grid) *n_grid = g.size();
  if (grid)
    for (size_t i = 0; i < g.size() && i < cap; ++i) grid[i] = g[i];
  if (n_dates) *n_dates = p.fixed_dates.size();
  if (dates)
    for (size_t i = 0; i < p.fixed_dates.size() && i < cap; ++i) dates[i] = p.fixed_dates[i];
  if (needs_fine) *needs_fine = p.needs_fine_grid ? 1 : 0;
  if (horizon) *horizon = p.horizon;
  if (assets && assets_cap) {
    std::string s;
    for (size_t i = 0; i < p.assets.size(); ++i) {
      if (i) s += '\n';
      s += p.assets[i];
    }
    snprintf(assets, assets_cap, "%s", s.c_str());
  }
  return KT_OK;
  KT_GUARD_END(nullptr)
}

kt_status kt_describe_program(const char* json, const kt_model* model, const kt_mc_config* cfg, char* buf,
                              uint64_t cap) {
  KT_GUARD_BEGIN
  ModelDesc md = model_from_abi(model);
  PreparedContract pc = prepare(json, md, cfg);
  Program prog = lower({pc.con.get()}, pc.grid, md.names);
  std::string s = prog.describe();
  if (buf && cap) snprintf(buf, cap, "%s", s.c_str());
  return KT_OK;
  KT_GUARD_END(nullptr)
}

}  // extern "C"
