// kt_kernels.hip — gfx950 path kernels: simulate, evaluate the contract's op
// program in registers, discount, and reduce in FP64, in one pass per path.
//
// Work layout: one path per lane, 256 lanes (4 waves) per workgroup, paths
// contiguous across the grid.  A path's state (log-spot per asset, Heston
// variance, the present values) stays in VGPRs; the op program's register
// file does too for one scenario and 16 registers, and lives (partly) in
// per-lane scratch for 32 registers and for the seven CRN scenarios of
// programs interpreted before their last two groups (the others interpret
// their tail one scenario at a time from a file in LDS).  Normals
// come from Philox4x32-10 keyed by the seed with counter
// (block, path_lo, path_hi, 0): normal j = k*M + c of grid point k, component
// c, is word j%4 of block j/4, so the stream does not depend on the launch
// shape.  Each group of 4 grid points consumes exactly M Philox blocks.
//
// Dispatch is per group of 4 grid points (KGroup): groups whose points all
// run one monitor (knock-out compare, running sum, running max/min) or none
// execute as unrolled straight-line code; the other groups interpret each
// point's op list.  The leading run of groups of one such kind (n_uniform)
// runs in a loop of its own before the dispatching loop takes over.  All
// tables are read through the constant address space, so their loads are
// scalar (s_load) and their values wave-uniform; the GBM and basket runs on
// grids of equal steps copy the shared step row into VGPRs instead.
//
// Scenarios (SC > 1, GBM only): SC repricings of one contract with different
// model parameters share every normal of a path (common random numbers); each
// scenario has its own log-spot, register file and present value, and the
// kernel reduces SC-weighted combinations of them (finite-difference Greeks).
#include <algorithm>

#include "../../include/kontracts.h"
#include "kt_path_ops.h"

namespace kt {

template <int MODEL, int K, int SC, bool DEBUG, class PA>
__device__ __forceinline__ void store_normals(const PA& A, int k, const float* zk, bool live, uint64_t local) {
  constexpr int M = MODEL == MK_GBM ? 1 : (MODEL == MK_HESTON ? 2 : K);
  if constexpr (DEBUG) {
    if (live && A.z_out) {
#pragma unroll
      for (int c = 0; c < M; ++c) A.z_out[((size_t)local * A.n_grid + k) * M + c] = zk[c];
    }
  }
}

// One interpreted grid point.
template <int MODEL, int K, int NS, int NOUT, int SC, bool DEBUG, class PA>
__device__ __forceinline__ void generic_point(const PA& A, int k, const float* zk, float (&x)[SC][K],
                                              float& v, float (&r)[SC][NS], double (&pv)[SC][NOUT], bool live,
                                              uint64_t local) {
  advance<MODEL, K, SC>(A, k, zk, x, v);
  store_normals<MODEL, K, SC, DEBUG>(A, k, zk, live, local);
  const uint64_t st = ld_u64(A.steps, (size_t)k);
  const int begin = (int)(uint32_t)st, count = (int)(uint32_t)(st >> 32);
  if (count > 0) run_ops<NS, NOUT, K, SC>(A, k, begin, count, r, pv, x, live);
}

// One group of 4 grid points under a single monitor, as straight-line code.
// The monitor's accumulators (acc, and comp for Kahan sums) are carried in
// registers across consecutive fast groups and only written back to the
// register file when the monitored slot changes or an interpreted group
// follows.  Knock-out groups keep the running max (min) of log-spot over the
// group in ext and compare it once with the threshold log2(H/s0): some point
// satisfies S >= H exactly when the group's maximum does.
template <int KIND>
constexpr bool knock_out_kind() {
  return KIND >= G_KO_GE && KIND <= G_KO_LT;
}

template <int KIND, int SC>
__device__ __forceinline__ void ko_init(float (&ext)[SC]) {
  constexpr bool UP = KIND == G_KO_GE || KIND == G_KO_GT;
  FOR_SC ext[sc] = UP ? -INFINITY : INFINITY;
}

// acc[sc] = 1 where scenario sc's extreme crosses its threshold thr[sc].
template <int KIND, int SC, class PA>
__device__ __forceinline__ void ko_apply(const PA& A, int g, const float (&ext)[SC], float (&acc)[SC]) {
  FOR_SC {
    const float thr = ld_f32(A.thr, (size_t)g * SC + sc);
    const float e = ext[xsrc<SC>(sc)];
    bool hit;
    if constexpr (KIND == G_KO_GE) hit = e >= thr;
    if constexpr (KIND == G_KO_GT) hit = e > thr;
    if constexpr (KIND == G_KO_LE) hit = e <= thr;
    if constexpr (KIND == G_KO_LT) hit = e < thr;
    acc[sc] = hit ? 1.0f : acc[sc];
  }
}

// The four points of a fast group: advance, and feed the monitor (knock-out
// kinds update ext only).  ROW: the step coefficients come from *R (VGPRs)
// instead of the table.
template <int KIND, int MODEL, int K, int SC, bool DEBUG, bool ROW = false, class PA>
__device__ __forceinline__ void fast_points(const PA& A, const KGroup gd, int k0, const float* z, float (&x)[SC][K],
                                            float& v, float (&acc)[SC], float (&comp)[SC], float (&ext)[SC],
                                            bool live, uint64_t local,
                                            const StepRow<MODEL, K, SC>* R = nullptr) {
  constexpr int M = MODEL == MK_GBM ? 1 : (MODEL == MK_HESTON ? 2 : K);
  constexpr bool KO = knock_out_kind<KIND>();
  constexpr bool UP = KIND == G_KO_GE || KIND == G_KO_GT;
  const uint32_t asset = (uint32_t)gd.asset < (uint32_t)K ? (uint32_t)gd.asset : 0u;
  // G_AVG_AT: the points that fix (wave-uniform)
  const uint32_t fix_mask = KIND == G_AVG_AT ? __float_as_uint(gd.imm) : 0xfu;
  float s0v[SC];
  FOR_SC s0v[sc] = scen_s0<K, SC>(A, sc, asset);
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    if constexpr (ROW) advance_row<MODEL, K, SC>(*R, &z[s * M], x, v);
    else advance<MODEL, K, SC>(A, k0 + s, &z[s * M], x, v);
    store_normals<MODEL, K, SC, DEBUG>(A, k0 + s, &z[s * M], live, local);
    if constexpr (KIND != G_NONE) {
      if (KIND == G_AVG_AT && !((fix_mask >> s) & 1u)) continue;
      FOR_SC {
        if (KO && xsrc<SC>(sc) != sc) continue;  // shares the extreme of its source scenario
        const float xs = K == 1 ? x[xsrc<SC>(sc)][0] : x[xsrc<SC>(sc)][asset];
        if constexpr (KO) ext[sc] = UP ? fmaxf(ext[sc], xs) : fminf(ext[sc], xs);
        if constexpr (KIND == G_AVG || KIND == G_AVG_AT || KIND == G_RMAX || KIND == G_RMIN) {
          const float S = s0v[sc] * exp2_fast(xs);
          if constexpr (KIND == G_AVG || KIND == G_AVG_AT) {
            const float y = S - comp[sc];
            const float t = acc[sc] + y;
            comp[sc] = (t - acc[sc]) - y;
            acc[sc] = t;
          }
          if constexpr (KIND == G_RMAX) acc[sc] = fmaxf(acc[sc], S);
          if constexpr (KIND == G_RMIN) acc[sc] = fminf(acc[sc], S);
        }
      }
    }
  }
}

template <int KIND, int MODEL, int K, int SC, bool DEBUG, class PA>
__device__ __forceinline__ void fast_group(const PA& A, const KGroup gd, int g, int k0, const float* z,
                                           float (&x)[SC][K], float& v, float (&acc)[SC], float (&comp)[SC],
                                           bool live, uint64_t local) {
  float ext[SC];
  if constexpr (knock_out_kind<KIND>()) ko_init<KIND, SC>(ext);
  fast_points<KIND, MODEL, K, SC, DEBUG>(A, gd, k0, z, x, v, acc, comp, ext, live, local);
  if constexpr (knock_out_kind<KIND>()) ko_apply<KIND, SC>(A, g, ext, acc);
}

// Fast kinds that get a loop of their own for a uniform prefix of groups.
template <int KIND, int MODEL>
constexpr bool uniform_loop_built() {
  if constexpr (MODEL == MK_BASKET) return KIND == G_NONE;
  return KIND == G_NONE || KIND == G_KO_GE || KIND == G_KO_LE || KIND == G_AVG || KIND == G_RMAX;
}

// Groups 0 .. n-1, all of kind KIND on gd's slot, asset and level: the same
// per-point code as the dispatching loop, in a loop of its own so that the
// monitor's accumulators stay in fixed registers across groups.  Knock-out
// groups of one level share their thresholds, so the extreme runs over the
// whole prefix and is compared once at its end: a point crosses the level
// exactly when the prefix's extreme does, and the flag only changes from
// its initial value to 1, so the result equals a compare per group.  On a
// grid of equal steps (tab_uniform) groups 1 .. n-1 take the step
// coefficients (and the Cholesky factor) from VGPR copies of table row 1
// (group 0 holds the zero row of point 0 and reads the table): 10% on the
// seven-scenario kernel, 2% on the 5-asset basket, 2-3% on the Heston
// European.  Heston runs under a running sum or extreme keep reading the
// table (7% slower with the copies: more VGPRs).
// UNROLL: groups per loop iteration (a tile knob: 4 or 8 grid points per
// iteration, the normals of both groups drawn before their steps; the same
// operations on the same values either way).  Written out by hand: the
// normal generator's wave ballot is convergent, which keeps the compiler
// from unrolling a loop with a run-time trip count.
template <int KIND, int MODEL, int K, int SC, bool DEBUG, int UNROLL, class PA>
__device__ __forceinline__ void uniform_groups(const PA& A, const KGroup gd, int n, uint32_t p_lo, uint32_t p_hi,
                                               uint32_t flip, float (&x)[SC][K], float& v, float (&acc)[SC],
                                               float (&comp)[SC], bool live, uint64_t local) {
  constexpr int M = MODEL == MK_GBM ? 1 : (MODEL == MK_HESTON ? 2 : K);
  float ext[SC];
  if constexpr (knock_out_kind<KIND>()) ko_init<KIND, SC>(ext);
  int g_table = 0;  // groups before g_table have run
  constexpr bool kRowKind = MODEL != MK_HESTON || !(KIND == G_AVG || KIND == G_RMAX || KIND == G_RMIN);
  if constexpr (kRowKind) {
    if (A.tab_uniform) {
      {
        float z[4 * M];
        group_normals<M>(0, p_lo, p_hi, A.seed_lo, A.seed_hi, flip, z);
        fast_points<KIND, MODEL, K, SC, DEBUG>(A, gd, 0, z, x, v, acc, comp, ext, live, local);
      }
      const StepRow<MODEL, K, SC> R = vgpr_row<MODEL, K, SC>(A, 1);
      int g = 1;
      if constexpr (UNROLL == 2) {
        // eight points per iteration: both groups' normals, then their steps
        for (; g + 1 < n; g += 2) {
          float z0[4 * M], z1[4 * M];
          group_normals<M>(g, p_lo, p_hi, A.seed_lo, A.seed_hi, flip, z0);
          group_normals<M>(g + 1, p_lo, p_hi, A.seed_lo, A.seed_hi, flip, z1);
          fast_points<KIND, MODEL, K, SC, DEBUG, true>(A, gd, 4 * g, z0, x, v, acc, comp, ext, live, local, &R);
          fast_points<KIND, MODEL, K, SC, DEBUG, true>(A, gd, 4 * g + 4, z1, x, v, acc, comp, ext, live, local, &R);
        }
      }
      for (; g < n; ++g) {
        float z[4 * M];
        group_normals<M>(g, p_lo, p_hi, A.seed_lo, A.seed_hi, flip, z);
        fast_points<KIND, MODEL, K, SC, DEBUG, true>(A, gd, 4 * g, z, x, v, acc, comp, ext, live, local, &R);
      }
      g_table = n;
    }
  }
  int g = g_table;
  if constexpr (UNROLL == 2) {
    for (; g + 1 < n; g += 2) {
      float z0[4 * M], z1[4 * M];
      group_normals<M>(g, p_lo, p_hi, A.seed_lo, A.seed_hi, flip, z0);
      group_normals<M>(g + 1, p_lo, p_hi, A.seed_lo, A.seed_hi, flip, z1);
      fast_points<KIND, MODEL, K, SC, DEBUG>(A, gd, 4 * g, z0, x, v, acc, comp, ext, live, local);
      fast_points<KIND, MODEL, K, SC, DEBUG>(A, gd, 4 * g + 4, z1, x, v, acc, comp, ext, live, local);
    }
  }
  for (; g < n; ++g) {
    float z[4 * M];
    group_normals<M>(g, p_lo, p_hi, A.seed_lo, A.seed_hi, flip, z);
    fast_points<KIND, MODEL, K, SC, DEBUG>(A, gd, 4 * g, z, x, v, acc, comp, ext, live, local);
  }
  if constexpr (knock_out_kind<KIND>()) ko_apply<KIND, SC>(A, 0, ext, acc);
}

// The fields run_ops reads, for one scenario of a scenario launch: its spot,
// its discount table and the launch's program (run_ops with SC = 1).
template <int K>
struct OneScenario {
  const KOp* ops;
  const double* df;
  int df_index[1];
  int n_grid;
  unsigned* err;
  float s0[K];
};

// One block of 256 paths: paths blk*256 .. blk*256+255 of the launch (global
// path index path_base + local), statistics into partials[blk].  Shared by
// the per-contract path kernel and the persistent portfolio kernel.
// r: the op program's register file, owned by the caller so that callers
// with several variants keep one register file for all of them.
// ANTI: each lane prices an antithetic pair (src/variance_reduction.rs:76-130):
// the path on the stream's normals, then the path on their negations, and
// keeps 0.5 * (pv_base + pv_anti) (src/variance_reduction.rs:196-201).
// Complementing every bit of a Philox word maps the uniform u of its cell to
// 1 - u, so normal_from_word(~w) == -normal_from_word(w) exactly.
// Tile knobs (the persistent portfolio kernel's variants): UNROLL groups per
// iteration of the uniform-run loops, and LANES lanes per workgroup: with
// LANES = 128 each lane prices two paths of the 256-path block one after the
// other (pass p takes paths p * 128 .. p * 128 + 127), and each wave's
// statistics land in the slot of the 64 paths it covered, so the block's
// partials are combined in the same order as with 256 lanes: the results do
// not depend on the tile.
template <int MODEL, int K, int NS, int NOUT, int SC, int NSTAT, bool DEBUG, bool ANTI = false, bool SERIAL = false,
          int UNROLL = 1, int LANES = kBlock, class PA>
__device__ __forceinline__ void path_block(const PA& A, uint32_t blk, float (&r)[SC][NS]) {
  static_assert(!ANTI || (SC == 1 && !DEBUG), "antithetic pairs: one scenario, no debug output");
  static_assert(!SERIAL || (SC > 1 && !DEBUG && MODEL == MK_GBM && K == 1), "scenario-serial tail: GBM scenarios");
  static_assert(LANES == kBlock || (LANES == kBlock / 2 && !SERIAL), "lanes per workgroup: 256, or 128 x 2 passes");
  constexpr int M = MODEL == MK_GBM ? 1 : (MODEL == MK_HESTON ? 2 : K);
  __shared__ KStats sh[NSTAT][kBlock / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll 1
  for (int pass = 0; pass < kBlock / LANES; ++pass) {
  const uint64_t local = (uint64_t)blk * kBlock + (uint64_t)pass * LANES + threadIdx.x;
  const bool live = local < A.n_paths;
  const uint64_t path = A.path_base + local;
  const uint32_t p_lo = (uint32_t)path, p_hi = (uint32_t)(path >> 32);

  double pv[SC][NOUT];
  double pair[NOUT];  // ANTI: the base path's pv while the antithetic one runs
  float x[SC][K];
#pragma unroll 1
  for (int rep = 0; rep < (ANTI ? 2 : 1); ++rep) {
  const uint32_t flip = (ANTI && rep) ? 0xffffffffu : 0u;
  FOR_SC {
#pragma unroll
    for (int j = 0; j < NOUT; ++j) pv[sc][j] = 0.0;
#pragma unroll
    for (int i = 0; i < K; ++i) x[sc][i] = 0.0f;
  }
  // The register file starts as zeros plus the init ops.  It is written only
  // where the dispatching loop below runs; the uniform prefix takes its
  // monitor's initial value from the init list directly.
  auto init_file = [&]() {
    FOR_SC {
#pragma unroll
      for (int i = 0; i < NS; ++i) r[sc][i] = 0.0f;
    }
    for (int i = 0; i < A.n_init; ++i) {
      const uint64_t w = ld_u64(A.init, (size_t)i);
      FOR_SC r[sc][((uint32_t)w >> 8) & (NS - 1)] = __uint_as_float((uint32_t)(w >> 32));
    }
  };
  auto init_value = [&](uint32_t slot) {
    float val = 0.0f;
    for (int i = 0; i < A.n_init; ++i) {
      const uint64_t w = ld_u64(A.init, (size_t)i);
      if ((((uint32_t)w >> 8) & (NS - 1)) == slot) val = __uint_as_float((uint32_t)(w >> 32));
    }
    return val;
  };
  float v = A.v0;
  // monitor accumulators cached in registers (slot mslot, kind mkind; -1 = none)
  float macc[SC], mcomp[SC];
  FOR_SC macc[sc] = mcomp[sc] = 0.0f;
  int mslot = -1, mkind = 0;
  // Slot access by compare-select chains: these run only at the boundaries of
  // monitor runs, and unlike indexed access they leave the register file in
  // place for the interpreter.
  auto put = [&](int sc, uint32_t slot, float val) {
#pragma unroll
    for (int i = 0; i < NS; ++i)
      if (slot == (uint32_t)i) r[sc][i] = val;
  };
  auto get = [&](int sc, uint32_t slot) {
    float val = r[sc][0];
#pragma unroll
    for (int i = 1; i < NS; ++i)
      if (slot == (uint32_t)i) val = r[sc][i];
    return val;
  };
  auto flush = [&]() {
    if (mslot >= 0) {
      const uint32_t s0i = (uint32_t)mslot & (NS - 1), s1i = (s0i + 1) & (NS - 1);
      FOR_SC {
        put(sc, s0i, macc[sc]);
        if (mkind == G_AVG || mkind == G_AVG_AT) put(sc, s1i, mcomp[sc]);
      }
      mslot = -1;
    }
  };

  auto load_group = [&](int g) {
    KGroup gd;
    gd.kind = (int32_t)ld_u32(A.groups, (size_t)g * 4);
    gd.slot = (int32_t)ld_u32(A.groups, (size_t)g * 4 + 1);
    gd.asset = (int32_t)ld_u32(A.groups, (size_t)g * 4 + 2);
    gd.imm = __uint_as_float(ld_u32(A.groups, (size_t)g * 4 + 3));
    return gd;
  };

  // A uniform prefix of groups runs in a loop specialised to its kind (the
  // same arithmetic in the same order as the dispatching loop below, which
  // then continues from where the prefix ends with the monitor still cached).
  int g_begin = 0;
  if constexpr (!DEBUG) {
    const int nu = min(A.n_uniform, A.n_groups);
    if (nu > 0) {
      const KGroup gu = load_group(0);
      auto enter = [&]() {
        if (gu.kind != G_NONE) {
          const uint32_t s0i = (uint32_t)gu.slot & (NS - 1), s1i = (s0i + 1) & (NS - 1);
          const float a0 = init_value(s0i), c0 = gu.kind == G_AVG ? init_value(s1i) : 0.0f;
          FOR_SC {
            macc[sc] = a0;
            mcomp[sc] = c0;
          }
          mslot = gu.slot;
          mkind = gu.kind;
        }
        g_begin = nu;
      };
      switch (gu.kind) {
#define KT_UNIFORM(KIND)                                                                              \
  case KIND:                                                                                          \
    if constexpr (uniform_loop_built<KIND, MODEL>()) {                                                \
      enter();                                                                                        \
      uniform_groups<KIND, MODEL, K, SC, DEBUG, UNROLL>(A, gu, nu, p_lo, p_hi, flip, x, v, macc, mcomp,  \
                                                        live, local);                                 \
    }                                                                                                 \
    break;
        KT_UNIFORM(G_NONE)
        KT_UNIFORM(G_KO_GE)
        KT_UNIFORM(G_KO_LE)
        KT_UNIFORM(G_AVG)
        KT_UNIFORM(G_RMAX)
#undef KT_UNIFORM
        default: break;
      }
    }
  }

  // Scenarios (SC > 1) after a uniform prefix that leaves at most two
  // groups: every scenario's register file then holds the init constants and
  // its monitor, so the remaining points are interpreted one scenario at a
  // time with one NS-register file (kept in VGPRs) instead of SC files
  // indexed at run time (which live in scratch).  Each scenario re-advances
  // its log-spot path (or its source scenario's) from the end of the prefix
  // over the same normals: the same operations on the same values, in the
  // same order per scenario, as the dispatching loop.
  // (SERIAL kernels only: the host launches them for programs whose uniform
  // prefix leaves at most two groups; any other program raises Unsupported)
  bool serial_done = false;
  if constexpr (SERIAL) {
    constexpr int TG = kSerialTailGroups;
    const int nt = A.n_groups - g_begin;
    if (!(g_begin > 0 && nt <= TG)) {
      if (live) atomicOr(A.err, 1u << KT_ERR_UNSUPPORTED);
    } else {
      float zt[TG][4];
#pragma unroll
      for (int t = 0; t < TG; ++t)
        if (t < nt) group_normals<1>(g_begin + t, p_lo, p_hi, A.seed_lo, A.seed_hi, flip, zt[t]);
      const uint32_t s0i = (uint32_t)mslot & (NS - 1), s1i = (s0i + 1) & (NS - 1);
      // One scenario per iteration (not unrolled: SC copies of the
      // interpreter would not fit the instruction cache); the scenario's
      // values are picked and written back by compare-selects, which keep
      // the per-scenario arrays in VGPRs.  The scenario's register file sits
      // in LDS, one column per lane (NS x 256 floats per workgroup): an op's
      // run-time register index becomes an LDS address instead of a chain of
      // selects over a VGPR array.
      __shared__ float lds_rf[NS * kBlock];
      const RegLds rf{&lds_rf[threadIdx.x], kBlock};
#pragma unroll 1
      for (int sc = 0; sc < SC; ++sc) {
        const int src = xsrc<SC>(sc);
        float m_acc = macc[0], m_comp = mcomp[0], x_start = x[0][0];
#pragma unroll
        for (int q = 1; q < SC; ++q) {
          if (sc == q) {
            m_acc = macc[q];
            m_comp = mcomp[q];
          }
          if (src == q) x_start = x[q][0];
        }
#pragma unroll
        for (int i = 0; i < NS; ++i) rf(0, (uint32_t)i) = 0.0f;
        for (int i = 0; i < A.n_init; ++i) {
          const uint64_t w = ld_u64(A.init, (size_t)i);
          rf(0, ((uint32_t)w >> 8) & (NS - 1)) = __uint_as_float((uint32_t)(w >> 32));
        }
        if (mslot >= 0) {
          rf(0, s0i) = m_acc;
          if (mkind == G_AVG || mkind == G_AVG_AT) rf(0, s1i) = m_comp;
        }
        float xs[1][1] = {{x_start}};
        double ps[1][NOUT];
#pragma unroll
        for (int j = 0; j < NOUT; ++j) ps[0][j] = 0.0;
        const OneScenario<1> V{A.ops, A.df, {A.df_index[sc]}, A.n_grid, A.err, {A.s0_scen[sc]}};
#pragma unroll
        for (int t = 0; t < TG; ++t) {
          if (t < nt) {
            const int k0 = 4 * (g_begin + t);
#pragma unroll
            for (int s = 0; s < 4; ++s) {
              const int k = k0 + s;
              if (k < A.n_grid) {
                const float a = ld_f32(A.tab, ((size_t)k * SC + src) * 2), b = ld_f32(A.tab, ((size_t)k * SC + src) * 2 + 1);
                xs[0][0] = fmaf(b, zt[t][s], xs[0][0] + a);
                const uint64_t st = ld_u64(A.steps, (size_t)k);
                const int begin = (int)(uint32_t)st, count = (int)(uint32_t)(st >> 32);
                if (count > 0) run_ops_rf<NS, NOUT, 1, 1>(V, k, begin, count, rf, ps, xs, live);
              }
            }
          }
        }
#pragma unroll
        for (int q = 0; q < SC; ++q)
          if (sc == q) {
#pragma unroll
            for (int j = 0; j < NOUT; ++j) pv[q][j] = ps[0][j];
          }
      }
      serial_done = true;
    }
  }
  if constexpr (!SERIAL) {
  init_file();
  for (int g = g_begin; g < A.n_groups; ++g) {
    // a later run of op-free groups (between the fixings of an average_at)
    // on a grid of equal steps: a loop of its own, from VGPR copies of row 1
    if constexpr (!DEBUG && MODEL != MK_BASKET && SC == 1) {
      if (A.tab_uniform && g > 0) {
        const int e = op_free_run_end(A, g);
        if (e > g) {
          const StepRow<MODEL, K, SC> R = vgpr_row<MODEL, K, SC>(A, 1);
          const KGroup none{G_NONE, 0, 0, 0.0f};
          float ext[SC];
          for (; g < e; ++g) {
            float z[4 * M];
            group_normals<M>(g, p_lo, p_hi, A.seed_lo, A.seed_hi, flip, z);
            fast_points<G_NONE, MODEL, K, SC, DEBUG, true>(A, none, 4 * g, z, x, v, macc, mcomp, ext, live, local, &R);
          }
        }
      }
    }
    const KGroup gd = load_group(g);
    float z[4 * M];
    group_normals<M>(g, p_lo, p_hi, A.seed_lo, A.seed_hi, flip, z);
    const int k0 = 4 * g;
    if (gd.kind != G_GENERIC && gd.kind != G_NONE && gd.slot != mslot) {
      flush();
      const uint32_t s0i = (uint32_t)gd.slot & (NS - 1), s1i = (s0i + 1) & (NS - 1);
      FOR_SC {
        macc[sc] = get(sc, s0i);
        mcomp[sc] = (gd.kind == G_AVG || gd.kind == G_AVG_AT) ? get(sc, s1i) : 0.0f;
      }
      mslot = gd.slot;
      mkind = gd.kind;
    }
    switch (gd.kind) {
#define KT_FAST(KIND) \
  case KIND: fast_group<KIND, MODEL, K, SC, DEBUG>(A, gd, g, k0, z, x, v, macc, mcomp, live, local); break;
      KT_FAST(G_NONE)
      KT_FAST(G_KO_GE)
      KT_FAST(G_KO_GT)
      KT_FAST(G_KO_LE)
      KT_FAST(G_KO_LT)
      KT_FAST(G_AVG)
      KT_FAST(G_AVG_AT)
      KT_FAST(G_RMAX)
      KT_FAST(G_RMIN)
#undef KT_FAST
      default: {
        flush();
        const int kend = min(4, A.n_grid - k0);
        if constexpr (MODEL == MK_BASKET) {
          // 4K normals exceed what indexed register access covers: unroll the
          // four points so every normal index is a compile-time constant.
#pragma unroll
          for (int s = 0; s < 4; ++s)
            if (s < kend)
              generic_point<MODEL, K, NS, NOUT, SC, DEBUG>(A, k0 + s, &z[s * M], x, v, r, pv, live, local);
        } else {
          for (int s = 0; s < kend; ++s) {
            // select this point's normals with uniform compares: a dynamic
            // index would move z out of registers for every group
            float zs[M];
#pragma unroll
            for (int c = 0; c < M; ++c) {
              float t = z[c];
#pragma unroll
              for (int q = 1; q < 4; ++q)
                if (s == q) t = z[q * M + c];
              zs[c] = t;
            }
            generic_point<MODEL, K, NS, NOUT, SC, DEBUG>(A, k0 + s, zs, x, v, r, pv, live, local);
          }
        }
      }
    }
  }
  flush();
  }  // !SERIAL
  (void)serial_done;
  if constexpr (ANTI) {
#pragma unroll
    for (int j = 0; j < NOUT; ++j) pair[j] = rep ? 0.5 * (pair[j] + pv[0][j]) : pv[0][j];
  }
  }  // rep
  if constexpr (ANTI) {
#pragma unroll
    for (int j = 0; j < NOUT; ++j) pv[0][j] = pair[j];
  }

  if constexpr (DEBUG) {
    if (live && A.pv_out) A.pv_out[local] = pv[0][0];
  }

  // Outputs: the pv accumulators (one scenario), or SC-weighted combinations.
  double y[NSTAT];
#pragma unroll
  for (int j = 0; j < NSTAT; ++j) {
    if constexpr (SC == 1) {
      y[j] = pv[0][j];
    } else {
      double t = 0.0;
      FOR_SC t += A.stat_w[j][sc] * pv[sc][0];
      y[j] = t;
    }
  }
  // One scenario, several contracts: output 0 may be the combination
  // sum_{o < combine} stat_w[0][o] pv_o of the first outputs (a control
  // variate's per-path pv_target - pv_control, src/variance_reduction.rs:226-231).
  if constexpr (SC == 1 && NOUT > 1) {
    if (A.combine > 0) {
      double t = 0.0;
#pragma unroll
      for (int o = 0; o < (NOUT < kScenarios ? NOUT : kScenarios); ++o)
        if (o < A.combine) t += A.stat_w[0][o] * pv[0][o];
      y[0] = t;
    }
  }
  // Block statistics per output: wave shifted sums (slot: the wave's 64
  // paths within the block), then Chan across the slots.
#pragma unroll
  for (int j = 0; j < NSTAT; ++j) {
    const KStats st = wave_stats(y[j], live);
    if (lane == 0) sh[j][pass * (LANES / 64) + wave] = st;
  }
  }  // pass
  __syncthreads();
  if (threadIdx.x < NSTAT) {
    KStats acc = sh[threadIdx.x][0];
#pragma unroll
    for (int wv = 1; wv < kBlock / 64; ++wv) acc = stats_combine(acc, sh[threadIdx.x][wv]);
    A.partials[(size_t)blk * NSTAT + threadIdx.x] = acc;
  }
  // sh is reused by the next block a persistent caller runs: no wave may
  // write it again before the reads above are done.
  __syncthreads();
}

template <int MODEL, int K, int NS, int NOUT, int SC, int NSTAT, bool DEBUG, bool ANTI, bool SERIAL = false>
__global__ __launch_bounds__(kBlock) void path_kernel(const PathArgs A) {
  float r[SC][NS];
  path_block<MODEL, K, NS, NOUT, SC, NSTAT, DEBUG, ANTI, SERIAL>(A, blockIdx.x, r);
}

// The seven-scenario CRN kernel allocates 123 VGPRs (4 waves per SIMD) for
// its payoff and reduction epilogue; capped at 102 (5 waves) the epilogue
// spills a few values and the path loop, which needs fewer, runs 3% faster
// (measured 1.006 against 1.038 ms per launch at 4,194,304 x 252; 6 waves
// cost 9% and 8 waves 29% through spills in the loop itself).  Programs of
// at most 8 registers use an 8-register file: the seven scenarios' files
// live in scratch, and halving them cut the kernel by 4% (0.87 against
// 0.907 ms on the same workload).
template <>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(5))) void
path_kernel<MK_GBM, 1, 16, 1, kScenarios, kScenarioStats, false, false>(const PathArgs A) {
  float r[kScenarios][16];
  path_block<MK_GBM, 1, 16, 1, kScenarios, kScenarioStats, false, false>(A, blockIdx.x, r);
}

template <>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(5))) void
path_kernel<MK_GBM, 1, 8, 1, kScenarios, kScenarioStats, false, false>(const PathArgs A) {
  float r[kScenarios][8];
  path_block<MK_GBM, 1, 8, 1, kScenarios, kScenarioStats, false, false>(A, blockIdx.x, r);
}

// The scenario-serial variants (programs whose uniform prefix leaves at most
// kSerialTailGroups groups, e.g. a knock-out monitored at every point and a
// payoff at the last): the tail's register file is one scenario's, in LDS,
// so the kernel has no register file in scratch: 0.85 against 0.89-0.93 ms
// on the 4,194,304 x 252 up-and-out, 2.4 MB of HBM writes per launch (the
// block partials) against 1.4 GB.  It needs 65 VGPRs (7 waves per SIMD);
// capped at 64 for 8 waves it runs at the same rate.
template <>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(5))) void
path_kernel<MK_GBM, 1, 16, 1, kScenarios, kScenarioStats, false, false, true>(const PathArgs A) {
  float r[kScenarios][16];
  path_block<MK_GBM, 1, 16, 1, kScenarios, kScenarioStats, false, false, true>(A, blockIdx.x, r);
}

template <>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(5))) void
path_kernel<MK_GBM, 1, 8, 1, kScenarios, kScenarioStats, false, false, true>(const PathArgs A) {
  float r[kScenarios][8];
  path_block<MK_GBM, 1, 8, 1, kScenarios, kScenarioStats, false, false, true>(A, blockIdx.x, r);
}

// Block j reduces output j.  err (nullable): block 0 copies the launch's
// error word to *err_copy and clears it for the next launch.
__global__ __launch_bounds__(256) void finalize_kernel(const KStats* __restrict__ partials, int n_blocks, int nout,
                                                       KStats* __restrict__ out, unsigned* err, unsigned* err_copy) {
  finalize_body(partials, n_blocks, nout, blockIdx.x, out + blockIdx.x);
  if (err && blockIdx.x == 0 && threadIdx.x == 0) {
    *err_copy = *err;
    *err = 0u;
  }
}

// ---------------------------------------------------------------------------
// Persistent portfolio kernel: a grid sized to the machine takes (job, block)
// work items from a queue ordered longest-first and runs each with the same
// path_block code as the per-contract launch, so every job's statistics are
// identical to pricing it alone.
// ---------------------------------------------------------------------------
// Clears the queue head and the jobs' error words (a kernel, not a memset, so
// that the launch sequence can be captured in a graph and replayed).
__global__ void portfolio_reset_kernel(unsigned* counter, unsigned* err, int n_jobs) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *counter = 0u;
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n_jobs; j += gridDim.x * blockDim.x) err[j] = 0u;
}

// WAVES > 0: the register budget for WAVES waves per SIMD (a tile the
// autotuner may pick); 0 leaves it to the compiler.
template <int NS, int WAVES, int UNROLL, int LANES>
__global__ __launch_bounds__(LANES) __attribute__((amdgpu_waves_per_eu(WAVES > 0 ? WAVES : 1))) void portfolio_kernel(
    const PortfolioArgs Q) {
  __shared__ uint32_t s_first;
  float r[1][NS];
  const KT_CONST PortfolioJob* jobs = (const KT_CONST PortfolioJob*)Q.jobs;
  const KT_CONST uint32_t* items = (const KT_CONST uint32_t*)Q.items;
  for (;;) {
    __syncthreads();  // every lane has read the previous s_first
    if (threadIdx.x == 0) s_first = atomicAdd(Q.counter, (unsigned)Q.chunk);
    __syncthreads();
    const uint32_t first = __builtin_amdgcn_readfirstlane(s_first);
    if (first >= Q.n_items) return;
    const uint32_t last = min(first + (uint32_t)Q.chunk, Q.n_items);
    for (uint32_t it = first; it < last; ++it) {
      const uint32_t j = items[2 * (size_t)it], blk = items[2 * (size_t)it + 1];
      const KT_CONST PathArgs& A = jobs[j].a;
      switch (jobs[j].variant) {
        case PV_GBM: path_block<MK_GBM, 1, NS, 1, 1, 1, false, false, false, UNROLL, LANES>(A, blk, r); break;
        case PV_HESTON: path_block<MK_HESTON, 1, NS, 1, 1, 1, false, false, false, UNROLL, LANES>(A, blk, r); break;
        case PV_BASKET2: path_block<MK_BASKET, 2, NS, 1, 1, 1, false, false, false, UNROLL, LANES>(A, blk, r); break;
        case PV_BASKET3: path_block<MK_BASKET, 3, NS, 1, 1, 1, false, false, false, UNROLL, LANES>(A, blk, r); break;
        case PV_BASKET4: path_block<MK_BASKET, 4, NS, 1, 1, 1, false, false, false, UNROLL, LANES>(A, blk, r); break;
        default: path_block<MK_BASKET, 5, NS, 1, 1, 1, false, false, false, UNROLL, LANES>(A, blk, r); break;
      }
    }
  }
}

// Block j finalizes job j's partials (the per-contract finalize, unchanged)
// and copies its error word to err_copy[j].
__global__ __launch_bounds__(256) void portfolio_finalize_kernel(const PortfolioJob* __restrict__ jobs,
                                                                 const int* __restrict__ n_blocks,
                                                                 KStats* __restrict__ out, unsigned* err_copy) {
  const int j = blockIdx.x;
  const KStats* partials = jobs[j].a.partials;
  finalize_body(partials, n_blocks[j], 1, 0, out + j);
  if (threadIdx.x == 0) err_copy[j] = *jobs[j].a.err;
}

// Greek surface: one seven-scenario CRN launch (the greeks_gbm kernel) per
// (spot, vol) cell, all cells in one 1-D grid of n_cells * n_blocks
// workgroups; each cell's PathArgs is read in place through the constant
// address space.
//
// Workgroup -> (cell, block) map: workgroups are dispatched round-robin over
// the XCDs and then over each XCD's CUs, so the workgroups resident on one CU
// at a time are L, L + n_cu, L + 2 n_cu, ...  Within each full chunk of
// n_cu * per_cu workgroups the map sends those co-resident workgroups to
// consecutive blocks of one cell, so a CU's scalar cache holds one cell's
// tables instead of per_cu of them.  The map is a bijection on every full
// chunk (identity on the last partial one), so it only changes which
// workgroup computes a block, never a result.
template <int NS, bool SERIAL>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(5))) void surface_kernel(
    const PathArgs* __restrict__ cells, uint32_t n_blocks, uint32_t total, uint32_t n_cu, uint32_t per_cu) {
  const uint32_t L = blockIdx.x;
  const uint32_t chunk = n_cu * per_cu;
  uint32_t v = L;
  const uint32_t q = L / chunk;
  if ((q + 1) * chunk <= total) {
    const uint32_t r = L - q * chunk;
    v = q * chunk + (r % n_cu) * per_cu + r / n_cu;
  }
  const uint32_t cell = v / n_blocks, blk = v - cell * n_blocks;
  const KT_CONST PathArgs& A = ((const KT_CONST PathArgs*)cells)[cell];
  float r[kScenarios][NS];
  path_block<MK_GBM, 1, NS, 1, kScenarios, kScenarioStats, false, false, SERIAL>(A, blk, r);
}

// Block c * kScenarioStats + j reduces output j of cell c (cell c's partials
// are [n_blocks][kScenarioStats] at partials + c * n_blocks * kScenarioStats).
__global__ __launch_bounds__(256) void surface_finalize_kernel(const KStats* __restrict__ partials, int n_blocks,
                                                               KStats* __restrict__ out) {
  const int cell = blockIdx.x / kScenarioStats, j = blockIdx.x % kScenarioStats;
  finalize_body(partials + (size_t)cell * n_blocks * kScenarioStats, n_blocks, kScenarioStats, j, out + blockIdx.x);
}

__global__ __launch_bounds__(256) void diff_stats_kernel(const double* __restrict__ a, const double* __restrict__ b,
                                                         double scale, uint64_t n, KStats* __restrict__ partials) {
  const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  const bool live = i < n;
  const double x = live ? (a[i] - b[i]) * scale : 0.0;
  const KStats st = wave_stats(x, live);
  __shared__ KStats sh[4];
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = st;
  __syncthreads();
  if (threadIdx.x == 0) {
    KStats t = sh[0];
#pragma unroll
    for (int w = 1; w < 4; ++w) t = stats_combine(t, sh[w]);
    partials[blockIdx.x] = t;
  }
}

__global__ void debug_philox_kernel(const uint32_t* __restrict__ ctr, uint32_t k0, uint32_t k1, uint64_t n,
                                    uint32_t* __restrict__ out) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint4 q = philox4x32_10(ctr[4 * i], ctr[4 * i + 1], ctr[4 * i + 2], ctr[4 * i + 3], k0, k1);
  out[4 * i] = q.x;
  out[4 * i + 1] = q.y;
  out[4 * i + 2] = q.z;
  out[4 * i + 3] = q.w;
}

// Plain streaming copy (16 B per lane per iteration): the known-size reference
// used to calibrate the HBM byte counters.
__global__ __launch_bounds__(256) void debug_copy_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                         uint64_t n16) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n16; i += (uint64_t)gridDim.x * 256)
    dst[i] = src[i];
}

// VALU issue calibration: eight independent FMA chains per lane, so each
// iteration issues eight VALU instructions and no chain waits on another;
// unrolled by 8 so that the scalar loop counter and branch come once per 64
// FMAs (iters is a multiple of 8).  The multiplier is held in a VGPR: with
// an SGPR operand the same loop issues at about half the rate.
__global__ __launch_bounds__(256) void debug_valu_kernel(float* __restrict__ out, int iters, float b) {
  float a[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) a[i] = (float)(threadIdx.x + i) * 1e-3f;
  const float bv = vgpr_copy(b);
#pragma unroll 8
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int i = 0; i < 8; ++i) a[i] = fmaf(a[i], bv, 1e-7f);
  }
  float s = 0.0f;
#pragma unroll
  for (int i = 0; i < 8; ++i) s += a[i];
  out[(size_t)blockIdx.x * 256 + threadIdx.x] = s;
}

// Normal-generator throughput: each lane draws `iters` Philox blocks with the
// path kernels' counter layout (block index, path, 0, 0; the key in SGPRs)
// and folds the four words (NORMALS false: one three-input xor and one xor
// per block) or their four inverse normals (NORMALS true: four adds per
// block) into one accumulator.  Gives the issue cost of the generator the
// path kernels run once per normal.
template <bool NORMALS>
__global__ __launch_bounds__(256) void debug_rng_rate_kernel(float* __restrict__ out, int iters, uint32_t k0,
                                                             uint32_t k1) {
  const uint32_t p = blockIdx.x * 256 + threadIdx.x;
  uint32_t h = 0;
  float s = 0.0f;
#pragma unroll 2
  for (int it = 0; it < iters; ++it) {
    const uint4 q = philox4x32_10((uint32_t)it, p, 0u, 0u, k0, k1);
    if constexpr (NORMALS)
      s += (normal_from_word(q.x) + normal_from_word(q.y)) + (normal_from_word(q.z) + normal_from_word(q.w));
    else
      h = xor3(h, q.x ^ q.y, xor3(q.z, q.w, 0u));
  }
  out[p] = NORMALS ? s : __uint_as_float(h);
}

__global__ void debug_normals_kernel(const uint32_t* __restrict__ w, uint64_t n, float* __restrict__ out) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = normal_from_word(w[i]);
}

// ---------------------------------------------------------------------------
// Longstaff-Schwartz (src/lsm.rs)
// ---------------------------------------------------------------------------
// Simulate each path on [0, dates...] and store, for every exercise date d,
// the exercised payoff value_at(d) and the regression spot (asset 0) as one
// FP32 pair, laid out [date][path] so that the backward passes stream them
// with coalesced 8-byte loads.  The payoff program has unit discount factors.
// VANILLA: the payoff at every date is scale * max(+-(S - strike), floor)
// (LsmPayoff), evaluated as straight-line code with the interpreter's
// operations in its order (spot, subtraction, max, the product with the
// init constant 1 and the scale), so the stored values are the same; the
// interpreter's per-op dispatch costs about 140 scalar instructions per
// date, which bound the kernel.
template <int MODEL, int K, int NS, bool VANILLA>
__global__ __launch_bounds__(kBlock) void lsm_paths_kernel(const PathArgs A, const LsmPayoff pay,
                                                           float2* __restrict__ cols, uint64_t stride) {
  constexpr int M = MODEL == MK_GBM ? 1 : (MODEL == MK_HESTON ? 2 : K);
  const uint64_t local = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  const bool live = local < A.n_paths;
  const uint64_t path = A.path_base + local;
  const uint32_t p_lo = (uint32_t)path, p_hi = (uint32_t)(path >> 32);
  float r[1][NS];
  float x[1][K];
#pragma unroll
  for (int i = 0; i < NS; ++i) r[0][i] = 0.0f;
#pragma unroll
  for (int i = 0; i < K; ++i) x[0][i] = 0.0f;
  for (int i = 0; i < A.n_init; ++i) {
    const uint64_t w = ld_u64(A.init, (size_t)i);
    r[0][((uint32_t)w >> 8) & (NS - 1)] = __uint_as_float((uint32_t)(w >> 32));
  }
  float v = A.v0;
  for (int g = 0; g < A.n_groups; ++g) {
    float z[4 * M];
#pragma unroll
    for (int bk = 0; bk < M; ++bk) {
      const uint4 q = philox4x32_10((uint32_t)(g * M + bk), p_lo, p_hi, 0u, A.seed_lo, A.seed_hi);
      z[4 * bk + 0] = normal_from_word(q.x);
      z[4 * bk + 1] = normal_from_word(q.y);
      z[4 * bk + 2] = normal_from_word(q.z);
      z[4 * bk + 3] = normal_from_word(q.w);
    }
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const int k = 4 * g + s;
      if (k < A.n_grid) {
        advance<MODEL, K, 1>(A, k, &z[s * M], x, v);
        if (k >= 1) {
          const float S = A.s0[0] * exp2_fast(x[0][0]);
          float payoff;
          if constexpr (VANILLA) {
            // no contraction of the spot's product into the subtraction: the
            // interpreter rounds the spot before it subtracts
#pragma clang fp contract(off)
            const float d = pay.kind == 1 ? S - pay.strike : pay.strike - S;
            payoff = (float)((double)(pay.scale * (fmaxf(d, pay.floor) * 1.0f)) * 1.0);
          } else {
            double pv[1][1] = {{0.0}};
            const uint64_t st = ld_u64(A.steps, (size_t)k);
            const int begin = (int)(uint32_t)st, count = (int)(uint32_t)(st >> 32);
            if (count > 0) run_ops<NS, 1, K, 1>(A, k, begin, count, r, pv, x, live);
            payoff = (float)pv[0][0];
          }
          if (live) cols[(size_t)(k - 1) * stride + local] = make_float2(payoff, S);
        }
      }
    }
  }
}

constexpr int kLsmMaxBasis = 6;
constexpr int kLsmPow = 2 * kLsmMaxBasis - 1;        // sums of S^p, p < 2m-1
constexpr int kLsmReg = kLsmPow + kLsmMaxBasis + 1;  // ... then S^j y, j < m, then the ITM count

struct LsmPass {
  const float2* cols;  // [n_dates][n] {payoff, spot}
  double* value;       // [n] realised value at the current date (t = 0 after the last pass)
  const double* coef;  // coefficients of date d_b: c_0..c_5, then 1 if fitted
  double* reg;         // [kLsmReg][gridDim.x] block partial sums of date d_a
  KStats* stats;       // [gridDim.x] final statistics
  uint64_t n;
  int n_dates;
  int d_b;  // apply the exercise decision at this date (-1: start from the last date's payoff)
  int d_a;  // accumulate the regression sums of this date (-1: none)
  int final_pass;
  double disc_b, disc_a, disc0;
};

// Sum of one value over the block's 256 threads in a fixed order, through
// LDS: each thread stores its value, 8 threads per value add 32 entries each,
// one thread adds the 8 partials.  red must hold NV x 257 doubles.
template <int NV>
__device__ __forceinline__ void block_sums(const double (&v)[NV], double* red, double* out) {
  constexpr int kRow = kBlock + 1;
#pragma unroll
  for (int j = 0; j < NV; ++j) red[j * kRow + threadIdx.x] = v[j];
  __syncthreads();
  double part = 0.0;
  const int j = threadIdx.x >> 3, seg = threadIdx.x & 7;
  if (j < NV) {
#pragma unroll 8
    for (int i = 0; i < kBlock / 8; ++i) part += red[j * kRow + seg * (kBlock / 8) + i];
  }
  __syncthreads();
  if (j < NV) red[j * 8 + seg] = part;
  __syncthreads();
  if (threadIdx.x < NV) {
    double t = 0.0;
#pragma unroll
    for (int q = 0; q < 8; ++q) t += red[threadIdx.x * 8 + q];
    out[threadIdx.x] = t;
  }
}

// One backward step, fused: apply the exercise decision of date d_b with its
// fitted coefficients (src/lsm.rs:203-222: an ITM path exercises iff its
// payoff beats the continuation estimate, the polynomial or, with no fit, the
// realised discounted value; the carried value is always realised), then
// accumulate date d_a's ITM normal equations (poly_fit, src/lsm.rs:259-288)
// in FP64 per block.  The last pass discounts to t = 0 and reduces the
// statistics.  Paths are visited grid-stride; all sums have a fixed order.
template <int MB>
__global__ __launch_bounds__(kBlock) void lsm_pass_kernel(const LsmPass P) {
  constexpr int NP = 2 * MB - 1;
  constexpr int NV = NP + MB + 1;
  const bool regress = P.d_a >= 0;
  double acc[NV];
#pragma unroll
  for (int j = 0; j < NV; ++j) acc[j] = 0.0;
  KStats st{0.0, 0.0, 0.0};
  bool have = false;
  double c[MB];
  if (P.d_b >= 0) {
    have = ld_f64(P.coef, kLsmMaxBasis) != 0.0;
#pragma unroll
    for (int j = 0; j < MB; ++j) c[j] = ld_f64(P.coef, j);
  }
  // Paths i = lane + t * step for trips t = 0, 1, ... (the same count on
  // every lane); four trips per iteration, their loads issued before any of
  // their arithmetic so that each lane keeps 12 loads in flight.  The sums
  // still take the paths in trip order.
  const uint64_t step = (uint64_t)gridDim.x * kBlock;
  const uint64_t trips = (P.n + step - 1) / step;
  const uint64_t lane0 = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  constexpr int U = 4;
  for (uint64_t t0 = 0; t0 < trips; t0 += U) {
    double vin[U];
    float2 pb[U], pa[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint64_t i = lane0 + (t0 + u) * step;
      const bool on = t0 + u < trips;  // uniform across the grid
      const uint64_t ii = on && i < P.n ? i : 0;
      vin[u] = 0.0;
      pb[u] = make_float2(0.0f, 0.0f);
      pa[u] = make_float2(0.0f, 0.0f);
      if (on) {
        if (P.d_b < 0) {
          pb[u] = P.cols[(size_t)(P.n_dates - 1) * P.n + ii];
        } else {
          vin[u] = P.value[ii];
          pb[u] = P.cols[(size_t)P.d_b * P.n + ii];
        }
        if (regress) pa[u] = P.cols[(size_t)P.d_a * P.n + ii];
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (t0 + u >= trips) break;
      const uint64_t i = lane0 + (t0 + u) * step;
      const bool live = i < P.n;
      const uint64_t ii = live ? i : 0;
      double val;
      if (P.d_b < 0) {
        val = (double)pb[u].x;
      } else {
        const double y = vin[u] * P.disc_b;
        const float2 ps = pb[u];
        val = y;
        if (ps.x > 0.0f) {
          double cont = y;
          if (have) {
            const double S = (double)ps.y;
            cont = c[MB - 1];
#pragma unroll
            for (int j = MB - 2; j >= 0; --j) cont = cont * S + c[j];
          }
          val = (double)ps.x > cont ? (double)ps.x : y;
        }
      }
      if (P.final_pass) {
        const double pv = val * P.disc0;
        if (live) P.value[ii] = pv;
        st = stats_combine(st, wave_stats(pv, live));
      } else if (live) {
        P.value[ii] = val;
      }
      if (regress) {
        const float2 ps = pa[u];
        if (live && ps.x > 0.0f) {
          const double S = (double)ps.y;
          const double y = val * P.disc_a;
          double pw = 1.0;
#pragma unroll
          for (int p = 0; p < NP; ++p) {
            acc[p] += pw;
            if (p < MB) acc[NP + p] = fma(pw, y, acc[NP + p]);
            pw *= S;
          }
          acc[NV - 1] += 1.0;
        }
      }
    }
  }
  __shared__ double red[kLsmReg * (kBlock + 1)];
  if (P.final_pass) {
    __shared__ KStats shs[kBlock / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) shs[wave] = st;
    __syncthreads();
    if (threadIdx.x == 0) {
      KStats t = shs[0];
#pragma unroll
      for (int wv = 1; wv < kBlock / 64; ++wv) t = stats_combine(t, shs[wv]);
      P.stats[blockIdx.x] = t;
    }
  }
  if (!regress) return;
  __shared__ double out[kLsmReg];
  block_sums<NV>(acc, red, out);
  __syncthreads();
  // layout [value][block]: sums S^p at 0..NP-1, S^j y at kLsmPow.., count last
  if (threadIdx.x < NV) {
    const int j = threadIdx.x;
    const int slot = j < NP ? j : (j < NP + MB ? kLsmPow + (j - NP) : kLsmReg - 1);
    P.reg[(size_t)slot * gridDim.x + blockIdx.x] = out[j];
  }
}

// Reduce the pass's block partials in a fixed order and solve the m x m
// normal equations by Gauss elimination with partial pivoting (solve_linear,
// src/numerics.rs:190-236: a pivot below 1e-12 gives zero coefficients).  No
// regression when fewer than m paths are ITM (src/lsm.rs:195).
template <int MB>
__global__ __launch_bounds__(kBlock) void lsm_solve_kernel(const double* __restrict__ reg, int n_blocks,
                                                           double* __restrict__ coef) {
  constexpr int NP = 2 * MB - 1;
  constexpr int NV = NP + MB + 1;
  // n_blocks <= kLsmPassBlocks: every load is independent and issues at once.
  double v[NV];
#pragma unroll
  for (int j = 0; j < NV; ++j) v[j] = 0.0;
#pragma unroll
  for (int q = 0; q < kLsmPassBlocks / kBlock; ++q) {
    const int b = threadIdx.x + q * kBlock;
    if (b < n_blocks) {
#pragma unroll
      for (int j = 0; j < NV; ++j) {
        const int slot = j < NP ? j : (j < NP + MB ? kLsmPow + (j - NP) : kLsmReg - 1);
        v[j] += reg[(size_t)slot * n_blocks + b];
      }
    }
  }
  __shared__ double red[kLsmReg * (kBlock + 1)];
  __shared__ double tot[kLsmReg];
  block_sums<NV>(v, red, tot);
  __syncthreads();
  if (threadIdx.x != 0) return;
  double a[MB][MB], b[MB], x[MB];
#pragma unroll
  for (int j = 0; j < MB; ++j) x[j] = 0.0;
  const bool have = tot[NV - 1] >= (double)MB;
  if (have) {
#pragma unroll
    for (int r0 = 0; r0 < MB; ++r0) {
#pragma unroll
      for (int c0 = 0; c0 < MB; ++c0) a[r0][c0] = tot[r0 + c0];
      b[r0] = tot[NP + r0];
    }
    bool singular = false;
#pragma unroll
    for (int col = 0; col < MB; ++col) {
      if (singular) break;
      // pivot: the first row of largest magnitude, then swap by selects so
      // that every index stays a compile-time constant
      int piv = col;
      double best = fabs(a[col][col]);
#pragma unroll
      for (int row = col + 1; row < MB; ++row) {
        const double m = fabs(a[row][col]);
        if (m > best) {
          best = m;
          piv = row;
        }
      }
      if (best < 1e-12) {
        singular = true;
        break;
      }
#pragma unroll
      for (int row = col + 1; row < MB; ++row) {
        const bool sw = row == piv;
#pragma unroll
        for (int c0 = 0; c0 < MB; ++c0) {
          const double t = a[col][c0];
          a[col][c0] = sw ? a[row][c0] : t;
          a[row][c0] = sw ? t : a[row][c0];
        }
        const double tb = b[col];
        b[col] = sw ? b[row] : tb;
        b[row] = sw ? tb : b[row];
      }
#pragma unroll
      for (int row = col + 1; row < MB; ++row) {
        const double f = a[row][col] / a[col][col];
#pragma unroll
        for (int c0 = col; c0 < MB; ++c0) a[row][c0] -= f * a[col][c0];
        b[row] -= f * b[col];
      }
    }
    if (!singular) {
#pragma unroll
      for (int row = MB - 1; row >= 0; --row) {
        double sum = b[row];
#pragma unroll
        for (int c0 = row + 1; c0 < MB; ++c0) sum -= a[row][c0] * x[c0];
        x[row] = sum / a[row][row];
      }
    }
  }
#pragma unroll
  for (int j = 0; j < kLsmMaxBasis; ++j) coef[j] = j < MB ? x[j] : 0.0;
  coef[kLsmMaxBasis] = have ? 1.0 : 0.0;
}

// ---------------------------------------------------------------------------
// Variant table and launchers.
// ---------------------------------------------------------------------------
namespace {

template <int MODEL, int K, int NS, int NOUT, int SC, int NSTAT, bool DEBUG, bool ANTI = false, bool SERIAL = false>
hipError_t go(const PathArgs& a, int n_blocks, hipStream_t s) {
  hipLaunchKernelGGL((path_kernel<MODEL, K, NS, NOUT, SC, NSTAT, DEBUG, ANTI, SERIAL>), dim3(n_blocks), dim3(kBlock), 0,
                     s, a);
  return hipGetLastError();
}

template <int MODEL, int K>
hipError_t go_ns(int ns, int nout, bool debug, bool anti, const PathArgs& a, int n_blocks, hipStream_t s) {
  if constexpr (MODEL == MK_GBM) {
    if (anti && !debug) {
      if (ns == 16 && nout == 1) return go<MODEL, K, 16, 1, 1, 1, false, true>(a, n_blocks, s);
      if (ns == 16 && nout == 8) return go<MODEL, K, 16, 8, 1, 8, false, true>(a, n_blocks, s);
      if (ns == 32 && nout == 1) return go<MODEL, K, 32, 1, 1, 1, false, true>(a, n_blocks, s);
      if (ns == 32 && nout == 8) return go<MODEL, K, 32, 8, 1, 8, false, true>(a, n_blocks, s);
      return hipErrorInvalidValue;
    }
  }
  if (anti) return hipErrorInvalidValue;
  if (debug) return go<MODEL, K, 32, 1, 1, 1, true>(a, n_blocks, s);
  if (ns == 16 && nout == 1) return go<MODEL, K, 16, 1, 1, 1, false>(a, n_blocks, s);
  if (ns == 16 && nout == 8) return go<MODEL, K, 16, 8, 1, 8, false>(a, n_blocks, s);
  if (ns == 32 && nout == 1) return go<MODEL, K, 32, 1, 1, 1, false>(a, n_blocks, s);
  if (ns == 32 && nout == 8) return go<MODEL, K, 32, 8, 1, 8, false>(a, n_blocks, s);
  return hipErrorInvalidValue;
}

}  // namespace

bool path_variant_exists(int model, int k_assets, int n_slots, int n_outputs, bool debug, int n_scen,
                         int* ns_chosen, int* nout_chosen, bool anti) {
  if (model == MK_BASKET ? (k_assets < 2 || k_assets > kMaxAssets) : k_assets != 1) return false;
  if (anti && (model != MK_GBM || n_scen != 1 || debug)) return false;
  if (n_scen != 1) {
    // CRN scenario kernel: one GBM contract, kScenarios scenarios, 8 or 16 registers.
    if (model != MK_GBM || n_scen != kScenarios || n_outputs != 1 || n_slots > 16 || debug) return false;
    *ns_chosen = n_slots <= 8 ? 8 : 16;
    *nout_chosen = 1;
    return true;
  }
  if (n_outputs < 1 || n_outputs > 8 || n_slots > 32) return false;
  int nout = n_outputs == 1 ? 1 : 8;
  int ns;
  if (debug) {
    if (n_outputs != 1) return false;
    ns = 32;
  } else {
    ns = n_slots <= 16 ? 16 : 32;
  }
  *ns_chosen = ns;
  *nout_chosen = nout;
  return true;
}

hipError_t launch_path(int model, int k, int ns, int nout, bool debug, int n_scen, const PathArgs& a, int n_blocks,
                       hipStream_t s, bool anti) {
  if (n_scen != 1) {
    if (anti) return hipErrorInvalidValue;
    const bool ser = a.serial_tail != 0;
    if (model == MK_GBM && n_scen == kScenarios && ns == 16 && nout == 1 && !debug)
      return ser ? go<MK_GBM, 1, 16, 1, kScenarios, kScenarioStats, false, false, true>(a, n_blocks, s)
                 : go<MK_GBM, 1, 16, 1, kScenarios, kScenarioStats, false>(a, n_blocks, s);
    if (model == MK_GBM && n_scen == kScenarios && ns == 8 && nout == 1 && !debug)
      return ser ? go<MK_GBM, 1, 8, 1, kScenarios, kScenarioStats, false, false, true>(a, n_blocks, s)
                 : go<MK_GBM, 1, 8, 1, kScenarios, kScenarioStats, false>(a, n_blocks, s);
    return hipErrorInvalidValue;
  }
  switch (model) {
    case MK_GBM: return go_ns<MK_GBM, 1>(ns, nout, debug, anti, a, n_blocks, s);
    case MK_HESTON: return go_ns<MK_HESTON, 1>(ns, nout, debug, anti, a, n_blocks, s);
    case MK_BASKET:
      switch (k) {
        case 2: return go_ns<MK_BASKET, 2>(ns, nout, debug, anti, a, n_blocks, s);
        case 3: return go_ns<MK_BASKET, 3>(ns, nout, debug, anti, a, n_blocks, s);
        case 4: return go_ns<MK_BASKET, 4>(ns, nout, debug, anti, a, n_blocks, s);
        case 5: return go_ns<MK_BASKET, 5>(ns, nout, debug, anti, a, n_blocks, s);
        case 6: return go_ns<MK_BASKET, 6>(ns, nout, debug, anti, a, n_blocks, s);
        case 7: return go_ns<MK_BASKET, 7>(ns, nout, debug, anti, a, n_blocks, s);
        case 8: return go_ns<MK_BASKET, 8>(ns, nout, debug, anti, a, n_blocks, s);
      }
  }
  return hipErrorInvalidValue;
}

hipError_t launch_lsm_paths(int model, int k, int ns, const PathArgs& a, const LsmPayoff& pay, int n_blocks,
                            float2* cols, uint64_t stride, hipStream_t s) {
#define KT_LSM(MODEL, KK, NSV) \
  hipLaunchKernelGGL((lsm_paths_kernel<MODEL, KK, NSV, false>), dim3(n_blocks), dim3(kBlock), 0, s, a, pay, cols, stride)
#define KT_LSM_V(MODEL) \
  hipLaunchKernelGGL((lsm_paths_kernel<MODEL, 1, 16, true>), dim3(n_blocks), dim3(kBlock), 0, s, a, pay, cols, stride)
  const bool vanilla = pay.kind == 1 || pay.kind == 2;
  if (model == MK_GBM) {
    if (vanilla) KT_LSM_V(MK_GBM);
    else if (ns == 16) KT_LSM(MK_GBM, 1, 16);
    else KT_LSM(MK_GBM, 1, 32);
  } else if (model == MK_HESTON) {
    if (vanilla) KT_LSM_V(MK_HESTON);
    else if (ns == 16) KT_LSM(MK_HESTON, 1, 16);
    else KT_LSM(MK_HESTON, 1, 32);
  } else {
    switch (k) {
      case 2: KT_LSM(MK_BASKET, 2, 32); break;
      case 3: KT_LSM(MK_BASKET, 3, 32); break;
      case 4: KT_LSM(MK_BASKET, 4, 32); break;
      case 5: KT_LSM(MK_BASKET, 5, 32); break;
      case 6: KT_LSM(MK_BASKET, 6, 32); break;
      case 7: KT_LSM(MK_BASKET, 7, 32); break;
      case 8: KT_LSM(MK_BASKET, 8, 32); break;
      default: return hipErrorInvalidValue;
    }
  }
#undef KT_LSM
#undef KT_LSM_V
  return hipGetLastError();
}

hipError_t launch_lsm_pass(const LsmPassArgs& p, int n_blocks, hipStream_t s) {
  LsmPass P{p.cols, p.value, p.coef, p.reg, p.stats, p.n, p.n_dates, p.d_b, p.d_a, p.final_pass,
            p.disc_b, p.disc_a, p.disc0};
  switch (p.m) {
    case 2: hipLaunchKernelGGL(lsm_pass_kernel<2>, dim3(n_blocks), dim3(kBlock), 0, s, P); break;
    case 3: hipLaunchKernelGGL(lsm_pass_kernel<3>, dim3(n_blocks), dim3(kBlock), 0, s, P); break;
    case 4: hipLaunchKernelGGL(lsm_pass_kernel<4>, dim3(n_blocks), dim3(kBlock), 0, s, P); break;
    case 5: hipLaunchKernelGGL(lsm_pass_kernel<5>, dim3(n_blocks), dim3(kBlock), 0, s, P); break;
    case 6: hipLaunchKernelGGL(lsm_pass_kernel<6>, dim3(n_blocks), dim3(kBlock), 0, s, P); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_lsm_solve(const double* reg, int n_blocks, int m, double* coef, hipStream_t s) {
  if (n_blocks < 1 || n_blocks > kLsmPassBlocks) return hipErrorInvalidValue;
  switch (m) {
    case 2: hipLaunchKernelGGL(lsm_solve_kernel<2>, dim3(1), dim3(kBlock), 0, s, reg, n_blocks, coef); break;
    case 3: hipLaunchKernelGGL(lsm_solve_kernel<3>, dim3(1), dim3(kBlock), 0, s, reg, n_blocks, coef); break;
    case 4: hipLaunchKernelGGL(lsm_solve_kernel<4>, dim3(1), dim3(kBlock), 0, s, reg, n_blocks, coef); break;
    case 5: hipLaunchKernelGGL(lsm_solve_kernel<5>, dim3(1), dim3(kBlock), 0, s, reg, n_blocks, coef); break;
    case 6: hipLaunchKernelGGL(lsm_solve_kernel<6>, dim3(1), dim3(kBlock), 0, s, reg, n_blocks, coef); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_diff_stats(const double* a, const double* b, double scale, uint64_t n, KStats* partials,
                             hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(diff_stats_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a, b, scale, n, partials);
  return hipGetLastError();
}

// The compiled tiles of the persistent price kernel: register budget (0:
// the compiler's, 6 or 8 waves per SIMD) x groups per loop iteration (1, 2)
// x lanes per workgroup (256, or 128 with two paths per lane).
#define KT_PF_TILES(X) \
  X(0, 1, 256) X(6, 1, 256) X(8, 1, 256) X(0, 2, 256) X(6, 2, 256) X(8, 2, 256) \
  X(0, 1, 128) X(6, 1, 128) X(8, 1, 128) X(0, 2, 128) X(6, 2, 128) X(8, 2, 128)

bool portfolio_tile_valid(const PortfolioTile& t) {
#define KT_PF_VALID(W, U, L) \
  if (t.waves == W && t.unroll == U && t.lanes == L) return true;
  KT_PF_TILES(KT_PF_VALID)
#undef KT_PF_VALID
  return false;
}

int portfolio_occupancy_blocks_per_cu(const PortfolioTile& t) {
  int n = 0;
  hipError_t e = hipErrorInvalidValue;
#define KT_PF_OCC(W, U, L)                                                                        \
  if (t.waves == W && t.unroll == U && t.lanes == L)                                              \
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, portfolio_kernel<16, W, U, L>, L, 0);
  KT_PF_TILES(KT_PF_OCC)
#undef KT_PF_OCC
  if (e != hipSuccess) return 4;
  return n > 0 ? n : 1;
}

hipError_t launch_portfolio(const PortfolioArgs& q, const int* n_blocks, int n_jobs, int grid, unsigned* err,
                            KStats* out, unsigned* err_copy, hipStream_t s, const PortfolioTile& t) {
  if (!portfolio_tile_valid(t)) return hipErrorInvalidValue;
  hipLaunchKernelGGL(portfolio_reset_kernel, dim3(std::max(1, std::min(64, (n_jobs + 255) / 256))), dim3(256), 0, s,
                     q.counter, err, n_jobs);
#define KT_PF_LAUNCH(W, U, L)                      \
  if (t.waves == W && t.unroll == U && t.lanes == L) \
    hipLaunchKernelGGL((portfolio_kernel<16, W, U, L>), dim3(grid), dim3(L), 0, s, q);
  KT_PF_TILES(KT_PF_LAUNCH)
#undef KT_PF_LAUNCH
  hipLaunchKernelGGL(portfolio_finalize_kernel, dim3(n_jobs), dim3(256), 0, s, q.jobs, n_blocks, out, err_copy);
  return hipGetLastError();
}
#undef KT_PF_TILES

int surface_occupancy_blocks_per_cu(int ns, bool serial) {
  int n = 0;
  hipError_t e;
  if (serial)
    e = ns == 8 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, surface_kernel<8, true>, kBlock, 0)
                : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, surface_kernel<16, true>, kBlock, 0);
  else
    e = ns == 8 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, surface_kernel<8, false>, kBlock, 0)
                : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, surface_kernel<16, false>, kBlock, 0);
  if (e != hipSuccess) return 4;
  return n > 0 ? n : 1;
}

hipError_t launch_surface(int ns, bool serial, const PathArgs* cells, int n_cells, int n_blocks,
                          const KStats* partials, KStats* out, int n_cu, int per_cu, hipStream_t s) {
  if (n_cells <= 0 || n_blocks <= 0 || n_cu <= 0 || per_cu <= 0 || (ns != 8 && ns != 16)) return hipErrorInvalidValue;
  const uint64_t total = (uint64_t)n_cells * (uint64_t)n_blocks;
  if (total > 0x7fffffffull || (uint64_t)n_cu * (uint64_t)per_cu > 0xffffffffull) return hipErrorInvalidValue;
#define KT_SURF(NSV, SER)                                                                                  \
  hipLaunchKernelGGL((surface_kernel<NSV, SER>), dim3((unsigned)total), dim3(kBlock), 0, s, cells, (uint32_t)n_blocks, \
                     (uint32_t)total, (uint32_t)n_cu, (uint32_t)per_cu)
  if (ns == 8) {
    if (serial) KT_SURF(8, true);
    else KT_SURF(8, false);
  } else {
    if (serial) KT_SURF(16, true);
    else KT_SURF(16, false);
  }
#undef KT_SURF
  hipLaunchKernelGGL(surface_finalize_kernel, dim3(n_cells * kScenarioStats), dim3(256), 0, s, partials, n_blocks,
                     out);
  return hipGetLastError();
}

hipError_t launch_finalize(const KStats* partials, int n_blocks, int nout, int n_used, KStats* out,
                           hipStream_t s, unsigned* err, unsigned* err_copy) {
  hipLaunchKernelGGL(finalize_kernel, dim3(n_used), dim3(256), 0, s, partials, n_blocks, nout, out, err, err_copy);
  return hipGetLastError();
}

hipError_t launch_debug_philox(const uint32_t* ctr, uint32_t k0, uint32_t k1, uint64_t n, uint32_t* out,
                               hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(debug_philox_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, ctr, k0, k1, n, out);
  return hipGetLastError();
}

hipError_t launch_debug_copy(const void* src, void* dst, uint64_t bytes, hipStream_t s) {
  const uint64_t n16 = bytes / 16;
  if (n16 == 0) return hipSuccess;
  hipLaunchKernelGGL(debug_copy_kernel, dim3(8192), dim3(256), 0, s, (const uint4*)src, (uint4*)dst, n16);
  return hipGetLastError();
}

hipError_t launch_debug_valu(float* out, int iters, hipStream_t s) {
  hipLaunchKernelGGL(debug_valu_kernel, dim3(kValuProbeBlocks), dim3(256), 0, s, out, iters, 1.0000001f);
  return hipGetLastError();
}

hipError_t launch_debug_rng_rate(bool normals, float* out, int iters, hipStream_t s) {
  if (normals)
    hipLaunchKernelGGL(debug_rng_rate_kernel<true>, dim3(kValuProbeBlocks), dim3(256), 0, s, out, iters, 42u, 7u);
  else
    hipLaunchKernelGGL(debug_rng_rate_kernel<false>, dim3(kValuProbeBlocks), dim3(256), 0, s, out, iters, 42u, 7u);
  return hipGetLastError();
}

hipError_t launch_debug_normals(const uint32_t* words, uint64_t n, float* out, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(debug_normals_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, words, n, out);
  return hipGetLastError();
}

}  // namespace kt
