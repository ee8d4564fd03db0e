// kt_portfolio.cpp — a heterogeneous book (different contracts, models,
// grids and path counts) priced in one persistent launch.  The reference
// prices such a book as a loop of price_gbm calls (src/pricer.rs:97-115);
// every job's result here is identical to kt_price on that job alone: same
// tables, same per-block path code, same per-job reduction order.
//
// Layout of the book in HBM (one upload): the job descriptors (PathArgs per
// job, pointing into the same buffer), the work items {job, block} ordered
// longest-first, each job's block count, then each job's program and model
// tables.  Each job owns a contiguous range of block partials and one error
// word.  The launch sequence (queue reset, persistent grid, per-job finalize)
// runs as three launches, or, with the engine's portfolio_graph option, is
// captured once in a hipGraph and replayed while the book's shape holds.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

#include "kt_engine_internal.h"

namespace kt {

namespace {

struct Blob {
  const void* p;
  size_t bytes;
};

struct JobPlan {
  std::shared_ptr<const Con> con;  // kept for variance-reduced jobs
  ModelDesc md;
  std::vector<double> grid;
  std::shared_ptr<const Program> prog;  // the engine's memoised lowering (shared, not copied)
  std::vector<KGroup> groups;
  std::vector<float> tab, thr;
  std::vector<double> df;
  int stride = 0;
  int variant = -1;
  uint64_t n_blocks = 0;
  double cost = 0.0;  // lane-ops per block, for the longest-first order
  size_t off[7] = {0, 0, 0, 0, 0, 0, 0};
};

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

int variant_of(const ModelDesc& md, const Program& prog) {
  if (prog.n_slots > 16 || prog.n_outputs != 1) return -1;
  if (md.kind == KT_MODEL_GBM) return PV_GBM;
  if (md.kind == KT_MODEL_HESTON) return PV_HESTON;
  const int K = (int)md.names.size();
  if (K >= 2 && K <= kPortfolioMaxAssets) return PV_BASKET2 + (K - 2);
  return -1;
}

}  // namespace

// A launched price-only book awaiting its stream: the plans, which jobs run
// in the persistent launch and which alone, and where the statistics land.
struct PriceBookRun {
  const kt_job* jobs = nullptr;
  uint64_t n_jobs = 0;
  kt_price_result* out = nullptr;
  const uint64_t* job_index = nullptr;
  int slot = 0;
  std::vector<JobPlan> plans;
  std::vector<uint64_t> solo;  // jobs priced by the per-contract path (after the stream syncs)
  std::vector<uint64_t> book;  // jobs in the persistent launch
  char* rb = nullptr;          // page-locked readback: statistics, then error words
};

void PriceBookRunDeleter::operator()(PriceBookRun* r) const { delete r; }

PriceBookHandle launch_price_book(kt_engine* E, const kt_job* jobs, uint64_t n_jobs, kt_price_result* out,
                                  const uint64_t* job_index, int slot) {
  PriceBookHandle run(new PriceBookRun());
  PriceBookRun& R = *run;
  R.jobs = jobs;
  R.n_jobs = n_jobs;
  R.out = out;
  R.job_index = job_index;
  R.slot = slot;
  // job numbers in messages: the caller's numbering when it passes a subset
  auto jid = [&](uint64_t j) { return std::to_string(job_index ? job_index[j] : j); };
  if (n_jobs == 0) return run;
  if (!jobs || !out) raise(KT_ERR_INVALID_ARGUMENT, "null jobs or output");
  if (n_jobs > (1ull << 24)) raise(KT_ERR_INVALID_ARGUMENT, "too many jobs");
  // Front half of every job, in order: the first failing job raises.
  std::vector<JobPlan>& plans = R.plans;
  plans.resize(n_jobs);
  std::vector<uint64_t>& solo = R.solo;
  std::vector<uint64_t>& book = R.book;
  for (uint64_t j = 0; j < n_jobs; ++j) {
    const kt_job& J = jobs[j];
    if (!J.contract_json) raise(KT_ERR_INVALID_ARGUMENT, "null contract in job " + jid(j));
    JobPlan& P = plans[j];
    P.md = model_from_abi(&J.model);
    PreparedContract pc = prepare(J.contract_json, P.md, &J.cfg, true, E);
    P.grid = std::move(pc.grid);
    out[j] = kt_price_result{};
    if (J.cfg.antithetic || J.cfg.control_variate) {
      // variance reduction runs its own launches, like a loop's price_gbm
      P.con = std::move(pc.con);
      solo.push_back(j);
      continue;
    }
    P.prog = lower_cached(E, pc.con, P.grid, P.md.names);
    P.variant = variant_of(P.md, *P.prog);
    if (J.cfg.n_paths == 0) {
      out[j] = result_from_stats(KStats{0.0, 0.0, 0.0});
      continue;
    }
    if (P.variant < 0) {
      solo.push_back(j);
      continue;
    }
    P.tab = model_table(P.md, P.grid, &P.stride);
    P.df.resize(P.grid.size());
    discount_factors(*P.prog, P.grid, J.cfg.rate, P.df.data());
    P.groups = knockout_tables(*P.prog, {P.md}, &P.thr);
    P.n_blocks = (J.cfg.n_paths + kBlock - 1) / kBlock;
    if (P.n_blocks > 0xffffffffull) raise(KT_ERR_INVALID_ARGUMENT, "too many paths in job " + jid(j));
    const double per_point = P.variant == PV_GBM ? 40.0 : P.variant == PV_HESTON ? 85.0 : 41.0 * P.md.names.size();
    P.cost = per_point * (double)P.grid.size() + 4.0 * (double)P.prog->ops.size();
    book.push_back(j);
  }
  if (book.empty()) return run;

  // Work items, longest job first (then by job index), blocks in order.
  std::vector<uint64_t> order = book;
  std::stable_sort(order.begin(), order.end(),
                   [&](uint64_t a, uint64_t b) { return plans[a].cost > plans[b].cost; });
  std::vector<uint32_t> pos(n_jobs, 0);  // job -> index in the launch
  for (size_t i = 0; i < book.size(); ++i) pos[book[i]] = (uint32_t)i;
  uint64_t n_items = 0, n_partials = 0;
  for (uint64_t j : book) n_items += plans[j].n_blocks;
  if (n_items > 0xffffffffull) raise(KT_ERR_INVALID_ARGUMENT, "book too large");
  std::vector<int> nblk(book.size());
  std::vector<uint64_t> part_off(book.size());
  for (size_t i = 0; i < book.size(); ++i) {
    nblk[i] = (int)plans[book[i]].n_blocks;
    part_off[i] = n_partials;
    n_partials += plans[book[i]].n_blocks;
  }
  // Arena layout: [jobs][items][nblk][tables of each job].
  const size_t jobs_bytes = align256(book.size() * sizeof(PortfolioJob));
  const size_t items_bytes = align256(2 * n_items * sizeof(uint32_t));
  const size_t nblk_bytes = align256(nblk.size() * sizeof(int));
  size_t cursor = jobs_bytes + items_bytes + nblk_bytes;
  for (size_t i = 0; i < book.size(); ++i) {
    JobPlan& P = plans[book[i]];
    const size_t sizes[7] = {std::max<size_t>(P.prog->ops.size(), 1) * sizeof(KOp), P.prog->steps.size() * sizeof(KStep),
                             P.groups.size() * sizeof(KGroup),   P.prog->init.size() * sizeof(KOp),
                             P.df.size() * sizeof(double),       P.tab.size() * sizeof(float),
                             P.thr.size() * sizeof(float)};
    for (int t = 0; t < 7; ++t) {
      P.off[t] = cursor;
      cursor += align256(std::max<size_t>(sizes[t], 1));
    }
  }
  DevBuf& book_buf = slot ? E->book2 : E->book;
  char* dev = (char*)E->ensure(book_buf, cursor);
  KStats* partials = (KStats*)E->ensure(slot ? E->partials2 : E->partials, n_partials * sizeof(KStats));
  // [statistics of each job][error word of each job]: one readback
  const size_t rb_bytes = book.size() * sizeof(KStats) + align256(book.size() * sizeof(unsigned));
  KStats* stats = (KStats*)E->ensure(slot ? E->stats2 : E->stats, std::max<size_t>(rb_bytes, kAdjMaxOut * sizeof(KStats)));
  unsigned* err_copy = (unsigned*)(stats + book.size());
  unsigned* errw = (unsigned*)E->ensure(slot ? E->errw2 : E->errw, align256((book.size() + 64) * sizeof(unsigned)));
  unsigned* counter = errw + book.size();

  // Everything the kernels read is written below; the alignment padding
  // between blobs is never read.
  char* H = E->stage(cursor, slot);
  {
    uint32_t* it = (uint32_t*)(H + jobs_bytes);
    for (uint64_t j : order)
      for (uint64_t b = 0; b < plans[j].n_blocks; ++b) {
        *it++ = pos[j];
        *it++ = (uint32_t)b;
      }
  }
  static const KOp kNop{OP_NOP, 0, 0, 0, 0.0f};
  for (size_t i = 0; i < book.size(); ++i) {
    JobPlan& P = plans[book[i]];
    const kt_job& J = jobs[book[i]];
    PortfolioJob pj{};
    PathArgs& A = pj.a;
    A.ops = (const KOp*)(dev + P.off[0]);
    A.steps = (const KStep*)(dev + P.off[1]);
    A.groups = (const KGroup*)(dev + P.off[2]);
    A.init = (const KOp*)(dev + P.off[3]);
    A.df = (const double*)(dev + P.off[4]);
    A.tab = (const float*)(dev + P.off[5]);
    A.thr = (const float*)(dev + P.off[6]);
    A.n_init = (int)P.prog->init.size();
    A.n_grid = (int)P.grid.size();
    A.n_groups = (int)P.groups.size();
    A.n_uniform = uniform_group_prefix(P.groups.data(), A.n_groups);
    A.tab_uniform = uniform_table_rows(P.tab, P.grid.size(), (size_t)P.stride);
    A.tab_stride = P.stride;
    A.n_paths = J.cfg.n_paths;
    A.path_base = 0;
    fill_model_args(A, P.md, J.cfg.seed);
    A.df_index[0] = 0;
    A.partials = partials + part_off[i];
    A.err = errw + i;
    pj.variant = P.variant;
    std::memcpy(H + i * sizeof(PortfolioJob), &pj, sizeof(PortfolioJob));
    const Blob blobs[7] = {{P.prog->ops.empty() ? &kNop : P.prog->ops.data(),
                            std::max<size_t>(P.prog->ops.size(), 1) * sizeof(KOp)},
                           {P.prog->steps.data(), P.prog->steps.size() * sizeof(KStep)},
                           {P.groups.data(), P.groups.size() * sizeof(KGroup)},
                           {P.prog->init.data(), P.prog->init.size() * sizeof(KOp)},
                           {P.df.data(), P.df.size() * sizeof(double)},
                           {P.tab.data(), P.tab.size() * sizeof(float)},
                           {P.thr.data(), P.thr.size() * sizeof(float)}};
    for (int t = 0; t < 7; ++t)
      if (blobs[t].bytes) std::memcpy(H + P.off[t], blobs[t].p, blobs[t].bytes);
  }
  std::memcpy(H + jobs_bytes + items_bytes, nblk.data(), nblk.size() * sizeof(int));
  E->stage_copy(dev, cursor, slot);

  PortfolioArgs Q{};
  Q.jobs = (const PortfolioJob*)dev;
  Q.items = (const uint32_t*)(dev + jobs_bytes);
  Q.n_items = (uint32_t)n_items;
  Q.counter = counter;
  const int* nblk_dev = (const int*)(dev + jobs_bytes + items_bytes);
  int n_cu = 0;
  check_hip(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, E->device), "device attribute");
  const PortfolioTile& tile = E->pf_tile;
  int& occ = E->occupancy_portfolio[tile.waves == 6 ? 1 : tile.waves == 8 ? 2 : 0][tile.unroll == 2 ? 1 : 0]
                                   [tile.lanes == kBlock ? 0 : 1];
  if (occ <= 0) occ = portfolio_occupancy_blocks_per_cu(tile);
  const int bpc = E->pf_blocks_per_cu > 0 ? E->pf_blocks_per_cu : occ;
  const uint64_t slots = (uint64_t)n_cu * bpc;
  // A queue chunk of c items saves atomics only where items outnumber the
  // resident workgroups c times over; on a smaller book it would serialise
  // items on a few workgroups.  Which workgroup takes an item never changes
  // a result.
  Q.chunk = (int)std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)std::max(1, E->pf_chunk), n_items / slots));
  const int grid = (int)std::max<uint64_t>(1, std::min<uint64_t>(slots, (n_items + Q.chunk - 1) / Q.chunk));

  // Replay the captured launch sequence while the book keeps its shape
  // (slot 0 only; slot 1 launches directly).
  const std::vector<uint64_t> key = {(uint64_t)dev, (uint64_t)partials, (uint64_t)stats, (uint64_t)errw,
                                     n_items, book.size(), (uint64_t)grid, (uint64_t)Q.chunk,
                                     (uint64_t)tile.waves, (uint64_t)tile.unroll, (uint64_t)tile.lanes};
  const bool graph = E->pf_use_graph && slot == 0;
  E->prof_begin();
  if (!graph) {
    check_hip(launch_portfolio(Q, nblk_dev, (int)book.size(), grid, errw, stats, err_copy, E->stream, tile),
              "portfolio launch");
  } else if (!E->pf_exec || E->pf_key != key) {
    if (E->pf_exec) (void)hipGraphExecDestroy(E->pf_exec);
    if (E->pf_graph) (void)hipGraphDestroy(E->pf_graph);
    E->pf_exec = nullptr;
    E->pf_graph = nullptr;
    E->pf_key.clear();
    check_hip(hipStreamBeginCapture(E->stream, hipStreamCaptureModeThreadLocal), "begin capture");
    hipError_t le =
        launch_portfolio(Q, nblk_dev, (int)book.size(), grid, errw, stats, err_copy, E->stream, tile);
    hipGraph_t g = nullptr;
    hipError_t ce = hipStreamEndCapture(E->stream, &g);
    check_hip(le, "portfolio launch");
    check_hip(ce, "end capture");
    E->pf_graph = g;
    check_hip(hipGraphInstantiate(&E->pf_exec, g, nullptr, nullptr, 0), "graph instantiate");
    E->pf_key = key;
  }
  if (graph) check_hip(hipGraphLaunch(E->pf_exec, E->stream), "graph launch");
  double path_steps = 0.0;
  for (uint64_t j : book) path_steps += (double)jobs[j].cfg.n_paths * (double)(plans[j].grid.size() - 1);
  E->prof_end(path_steps);

  R.rb = E->readback(rb_bytes, slot);
  check_hip(hipMemcpyAsync(R.rb, stats, rb_bytes, hipMemcpyDeviceToHost, E->stream), "copy stats");
  return run;
}

void finish_price_book(kt_engine* E, PriceBookRun& R) {
  if (R.n_jobs == 0) return;
  auto jid = [&](uint64_t j) { return std::to_string(R.job_index ? R.job_index[j] : j); };
  auto price_solo = [&](uint64_t j) {
    const JobPlan& P = R.plans[j];
    const kt_mc_config& c = R.jobs[j].cfg;
    if (P.con)
      return price_vr(E, *P.con, P.md, P.grid, c.rate, c.n_paths, c.seed, c.antithetic != 0, c.control_variate != 0);
    return result_from_stats(run_program(E, *P.prog, P.md, P.grid, c.rate, c.n_paths, c.seed, nullptr, nullptr)[0]);
  };
  if (R.book.empty()) {
    for (uint64_t j : R.solo) R.out[j] = price_solo(j);
    return;
  }
  check_hip(hipStreamSynchronize(E->stream), "stream sync");
  const size_t nb = R.book.size();
  std::vector<KStats> st(nb);
  std::vector<unsigned> err_bits(nb);
  std::memcpy(st.data(), R.rb, nb * sizeof(KStats));
  std::memcpy(err_bits.data(), R.rb + nb * sizeof(KStats), nb * sizeof(unsigned));
  // A path error fails the call like the reference loop would at that job,
  // unless an earlier job (solo or in the book) fails first.
  uint64_t first_err = R.n_jobs;
  int first_code = 0;
  for (size_t i = 0; i < nb; ++i)
    if (err_bits[i] && R.book[i] < first_err) {
      first_err = R.book[i];
      first_code = lowest_error_code(err_bits[i]);
    }
  for (uint64_t j : R.solo) {
    if (j > first_err) break;
    R.out[j] = price_solo(j);
  }
  if (first_code)
    raise(first_code, error_text(first_code, *R.plans[first_err].prog) + std::string(" (job ") + jid(first_err) + ")");
  for (size_t i = 0; i < nb; ++i) R.out[R.book[i]] = result_from_stats(st[i]);
}

void price_portfolio(kt_engine* E, const kt_job* jobs, uint64_t n_jobs, kt_price_result* out,
                     const uint64_t* job_index) {
  PriceBookHandle run = launch_price_book(E, jobs, n_jobs, out, job_index, 0);
  finish_price_book(E, *run);
}

}  // namespace kt
