// kt_portfolio_greeks.cpp — a heterogeneous book priced together with the
// pathwise Greeks of every job whose payoff admits them (one flow date,
// arithmetic ops, running sums / extremes fed by a spot: European, Asian,
// lookback and basket payoffs).  Those jobs run in persistent launches over
// (job, block) queues, one launch per kernel variant (GBM, Heston, Heston with
// running accumulators, 5-asset basket), each job's result equal to
// kt_greeks_adjoint on that job alone.  The other jobs (barriers, digitals,
// programs beyond the variants' registers, variance-reduced configurations)
// are priced by the persistent price-only launch without Greeks, or, when no
// portfolio variant covers their model, by a per-contract Greeks launch.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <numeric>

#include "kt_engine_internal.h"

namespace kt {

namespace {

size_t a256(size_t b) { return (b + 255) & ~(size_t)255; }

struct AdjJob {
  uint64_t index = 0;  // position in the caller's book
  ModelDesc md;
  std::vector<double> grid;
  std::shared_ptr<const Program> prog;  // the engine's memoised lowering (shared, not copied)
  AdjPlan plan;
  double rate = 0.0;
  AdjJobHost H;
  int variant = -1;
  int need = 0;
  uint64_t n_blocks = 0;
  double cost = 0.0;
  size_t off[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  uint64_t part_off = 0;
};

}  // namespace

void price_portfolio_greeks(kt_engine* E, const kt_job* jobs, uint64_t n_jobs, kt_price_result* out, double* greeks,
                            double* greeks_se, uint64_t stride) {
  if (n_jobs == 0) return;
  if (!jobs || !out) raise(KT_ERR_INVALID_ARGUMENT, "null jobs or output");
  if (n_jobs > (1ull << 24)) raise(KT_ERR_INVALID_ARGUMENT, "too many jobs");
  if ((greeks || greeks_se) && stride < (uint64_t)(kAdjMaxOut - 1))
    raise(KT_ERR_INVALID_ARGUMENT, "Greek rows need a stride of at least " + std::to_string(kAdjMaxOut - 1));
  const double nan = std::numeric_limits<double>::quiet_NaN();
  for (uint64_t j = 0; j < n_jobs; ++j) {
    out[j] = kt_price_result{};
    for (uint64_t p = 0; p < stride; ++p) {
      if (greeks) greeks[j * stride + p] = nan;
      if (greeks_se) greeks_se[j * stride + p] = nan;
    }
  }
  auto put_greeks = [&](uint64_t j, const std::vector<KStats>& st, size_t first, int need) {
    out[j] = result_from_stats(st[first]);
    for (int p = 0; p + 1 < need; ++p) {
      const kt_price_result r = result_from_stats(st[first + 1 + p]);
      if (greeks) greeks[j * stride + p] = r.price;
      if (greeks_se) greeks_se[j * stride + p] = r.std_error;
    }
  };

  // Front half of every job, in order: the first failing job raises.
  std::vector<uint64_t> price_only, solo;
  std::vector<AdjJob> adj;
  for (uint64_t j = 0; j < n_jobs; ++j) {
    const kt_job& J = jobs[j];
    if (!J.contract_json) raise(KT_ERR_INVALID_ARGUMENT, "null contract in job " + std::to_string(j));
    ModelDesc md = model_from_abi(&J.model);
    const bool vr = J.cfg.antithetic || J.cfg.control_variate;
    PreparedContract pc = prepare(J.contract_json, md, &J.cfg, vr, E);
    std::shared_ptr<const Program> prog_ptr = lower_cached(E, pc.con, pc.grid, md.names);
    const Program& prog = *prog_ptr;
    bool ok = !vr;
    for (const KOp& op : prog.ops) ok = ok && op.code != OP_FLAGERR;  // path errors: the pricer raises them
    AdjPlan plan;
    if (ok) {
      const std::shared_ptr<const AdjPlan> cached = adjoint_plan_cached(E, prog_ptr);
      if (cached) plan = *cached;
      else ok = false;
    }
    const int kind = model_kernel_kind(md), K = (int)md.names.size();
    int ns = 0;
    if (ok) ok = adjoint_variant_exists(kind, K, prog.n_slots, &ns);
    if (!ok) {
      price_only.push_back(j);
      continue;
    }
    const int need = adjoint_outputs(kind, K);
    if (J.cfg.n_paths == 0) {  // as kt_greeks_adjoint: zero price and derivatives
      for (int p = 0; p + 1 < need; ++p) {
        if (greeks) greeks[j * stride + p] = 0.0;
        if (greeks_se) greeks_se[j * stride + p] = 0.0;
      }
      continue;
    }
    const int variant = portfolio_adjoint_variant(kind, K, prog.n_slots, plan.n_acc, E->pathwise_tangent);
    if (variant < 0) {
      solo.push_back(j);
      continue;
    }
    AdjJob a;
    a.index = j;
    a.md = std::move(md);
    a.grid = std::move(pc.grid);
    a.prog = std::move(prog_ptr);
    a.plan = std::move(plan);
    a.rate = J.cfg.rate;
    a.variant = variant;
    a.need = need;
    a.n_blocks = (J.cfg.n_paths + kBlock - 1) / kBlock;
    if (a.n_blocks > 0xffffffffull) raise(KT_ERR_INVALID_ARGUMENT, "too many paths in job " + std::to_string(j));
    const double per_point = variant == APV_GBM ? 60.0 : variant == APV_BASKET5 ? 260.0 : 250.0;
    a.cost = per_point * (double)a.grid.size() + 4.0 * (double)a.prog->ops.size();
    adj.push_back(std::move(a));
  }

  // The price-only jobs go first, in the second staging slot and buffer
  // set: their persistent launch runs while the Greeks launches below are
  // prepared on the host.
  std::vector<kt_job> po_jobs;
  std::vector<kt_price_result> po_out(price_only.size());
  PriceBookHandle po_run;
  if (!price_only.empty()) {
    po_jobs.reserve(price_only.size());
    for (uint64_t j : price_only) po_jobs.push_back(jobs[j]);
    po_run = launch_price_book(E, po_jobs.data(), po_jobs.size(), po_out.data(), price_only.data(), 1);
  }
  for (AdjJob& a : adj) a.H = adjoint_job_host(*a.prog, a.md, a.grid, a.rate, std::move(a.plan));

  if (!adj.empty()) {
    const size_t n_adj = adj.size();
    // Items per variant, longest job first (then by book order), blocks in order.
    std::vector<std::vector<uint32_t>> vitems(APV_COUNT);
    std::vector<size_t> order(n_adj);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return adj[a].cost > adj[b].cost; });
    uint64_t n_items = 0, n_partials = 0;
    for (size_t i : order) {
      for (uint64_t b = 0; b < adj[i].n_blocks; ++b) {
        vitems[adj[i].variant].push_back((uint32_t)i);
        vitems[adj[i].variant].push_back((uint32_t)b);
      }
      n_items += adj[i].n_blocks;
    }
    if (n_items > 0xffffffffull) raise(KT_ERR_INVALID_ARGUMENT, "book too large");
    for (AdjJob& a : adj) {
      a.part_off = n_partials;
      n_partials += a.n_blocks * (uint64_t)a.need;
    }
    // Arena: [AdjArgs per job][items of each variant][n_blocks][n_out][per job: 10 table blobs]
    size_t cursor = a256(n_adj * sizeof(AdjArgs));
    size_t items_off[APV_COUNT];
    for (int v = 0; v < APV_COUNT; ++v) {
      items_off[v] = cursor;
      cursor += a256(std::max<size_t>(vitems[v].size(), 1) * sizeof(uint32_t));
    }
    const size_t nblk_off = cursor;
    cursor += a256(n_adj * sizeof(int));
    const size_t nout_off = cursor;
    cursor += a256(n_adj * sizeof(int));
    std::vector<std::vector<std::pair<const void*, size_t>>> blobs(n_adj);
    for (size_t i = 0; i < n_adj; ++i) {
      blobs[i] = adjoint_blobs(adj[i].H, *adj[i].prog);
      for (int t = 0; t < 10; ++t) {
        adj[i].off[t] = cursor;
        cursor += a256(std::max<size_t>(blobs[i][t].second, 1));
      }
    }
    char* dev = (char*)E->ensure(E->book, cursor);
    KStats* partials = (KStats*)E->ensure(E->partials, n_partials * sizeof(KStats));
    KStats* stats = (KStats*)E->ensure(E->stats, n_adj * kAdjMaxOut * sizeof(KStats));
    unsigned* counters = (unsigned*)E->ensure(E->errw, 256);
    // Heston checkpoint slots: one per resident workgroup of the larger launch.
    int n_cu = 0;
    check_hip(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, E->device), "device attribute");
    int grid_v[APV_COUNT] = {0, 0, 0, 0, 0, 0}, chunk_v[APV_COUNT] = {1, 1, 1, 1, 1, 1};
    for (int v = 0; v < APV_COUNT; ++v) {
      if (vitems[v].empty()) continue;
      int& occ = E->occupancy_adjoint[E->pf_adj_waves == 5 ? 1 : 0][v];
      if (occ <= 0) occ = portfolio_adjoint_occupancy(v, E->pf_adj_waves);
      const int bpc = E->pf_blocks_per_cu > 0 ? std::min(E->pf_blocks_per_cu, occ) : occ;
      // as for prices: no chunk larger than the items per resident workgroup
      const uint64_t slots = (uint64_t)n_cu * bpc, n_it = vitems[v].size() / 2;
      chunk_v[v] = (int)std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)std::max(1, E->pf_chunk), n_it / slots));
      grid_v[v] = (int)std::max<uint64_t>(1, std::min<uint64_t>(slots, (n_it + chunk_v[v] - 1) / chunk_v[v]));
    }
    uint64_t ck_floats = 0;
    for (const AdjJob& a : adj)
      if (a.variant == APV_HESTON || a.variant == APV_HESTON_FEEDS)
        ck_floats = std::max<uint64_t>(ck_floats, ((a.grid.size() - 1) / kAdjSeg + 1) * 2 * kBlock);
    float* ck = nullptr;
    if (ck_floats)
      ck = (float*)E->ensure(E->aux,
                             (size_t)std::max(grid_v[APV_HESTON], grid_v[APV_HESTON_FEEDS]) * ck_floats * sizeof(float));

    // every byte the kernels read is written below (padding is never read)
    char* H = E->stage(cursor);
    std::vector<int> nblk(n_adj), nout(n_adj);
    for (size_t i = 0; i < n_adj; ++i) {
      AdjJob& a = adj[i];
      const kt_job& J = jobs[a.index];
      std::vector<void*> dptr(10);
      for (int t = 0; t < 10; ++t) {
        dptr[t] = dev + a.off[t];
        if (blobs[i][t].second) std::memcpy(H + a.off[t], blobs[i][t].first, blobs[i][t].second);
      }
      AdjArgs A{};
      fill_adjoint_args(A, a.H, *a.prog, a.md, a.grid, J.cfg.seed, dptr);
      A.p.n_paths = J.cfg.n_paths;
      A.p.path_base = 0;
      A.p.partials = partials + a.part_off;
      A.p.err = counters + APV_COUNT;  // never written: the covered programs raise no path errors
      A.seg = kAdjSeg;
      A.tangent = (a.variant == APV_HESTON_T || a.variant == APV_HESTON_FEEDS_T) ? 1 : 0;
      std::memcpy(H + i * sizeof(AdjArgs), &A, sizeof(AdjArgs));
      nblk[i] = (int)a.n_blocks;
      nout[i] = a.need;
    }
    for (int v = 0; v < APV_COUNT; ++v)
      if (!vitems[v].empty()) std::memcpy(H + items_off[v], vitems[v].data(), vitems[v].size() * sizeof(uint32_t));
    std::memcpy(H + nblk_off, nblk.data(), n_adj * sizeof(int));
    std::memcpy(H + nout_off, nout.data(), n_adj * sizeof(int));
    E->stage_copy(dev, cursor);
    check_hip(hipMemsetAsync(counters, 0, (APV_COUNT + 1) * sizeof(unsigned), E->stream), "memset");

    E->prof_begin();
    for (int v = 0; v < APV_COUNT; ++v) {
      if (vitems[v].empty()) continue;
      AdjPortfolioArgs Q{};
      Q.jobs = (const AdjArgs*)dev;
      Q.items = (const uint32_t*)(dev + items_off[v]);
      Q.n_items = (uint32_t)(vitems[v].size() / 2);
      Q.chunk = chunk_v[v];
      Q.counter = counters + v;
      Q.ck = (v == APV_HESTON || v == APV_HESTON_FEEDS) ? ck : nullptr;
      Q.ck_floats = ck_floats;
      check_hip(launch_portfolio_adjoint(v, Q, grid_v[v], E->stream, E->pf_adj_waves), "portfolio adjoint launch");
    }
    double path_steps = 0.0;
    for (const AdjJob& a : adj) path_steps += (double)jobs[a.index].cfg.n_paths * (double)(a.grid.size() - 1);
    E->prof_end(path_steps);
    check_hip(launch_portfolio_adjoint_finalize((const AdjArgs*)dev, (const int*)(dev + nblk_off),
                                                (const int*)(dev + nout_off), (int)n_adj, stats, E->stream),
              "portfolio adjoint finalize");
    const size_t rb_bytes = n_adj * kAdjMaxOut * sizeof(KStats);
    char* rb = E->readback(rb_bytes);
    check_hip(hipMemcpyAsync(rb, stats, rb_bytes, hipMemcpyDeviceToHost, E->stream), "copy stats");
    check_hip(hipStreamSynchronize(E->stream), "stream sync");
    std::vector<KStats> st(n_adj * kAdjMaxOut);
    std::memcpy(st.data(), rb, rb_bytes);
    for (size_t i = 0; i < n_adj; ++i) put_greeks(adj[i].index, st, i * kAdjMaxOut, adj[i].need);
  }

  // Jobs without Greeks: the persistent price-only launch (errors numbered in
  // the caller's book).
  if (po_run) {
    finish_price_book(E, *po_run);
    for (size_t i = 0; i < price_only.size(); ++i) out[price_only[i]] = po_out[i];
  }
  // Greeks jobs no portfolio variant covers: one launch each.
  for (uint64_t j : solo) {
    std::vector<double> v(kAdjMaxOut), se(kAdjMaxOut);
    std::vector<KStats> st;
    greeks_adjoint(E, jobs[j].contract_json, &jobs[j].model, &jobs[j].cfg, v.data(), se.data(), kAdjMaxOut, &st);
    put_greeks(j, st, 0, (int)st.size());
  }
}

}  // namespace kt
