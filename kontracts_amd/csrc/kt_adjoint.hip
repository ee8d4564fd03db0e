// kt_adjoint.hip — pathwise Greeks on gfx950 (no reference counterpart: the
// reference has bump-and-reprice Greeks only, src/greeks.rs).
//
// Supported payoffs pay on one date, the last grid point n, an arithmetic
// function of the spots at n and of running sums / extremes of spots.  The
// per-path sample of d(price)/d(theta) is the pathwise derivative
//   sum over leaves L of  d(pv)/dL * dL/dtheta,
// where the leaves are the spots at n and the accumulators.  d(pv)/dL comes
// from a dual-number pass over point n's ops; dL/dtheta from the dynamics:
//
// * Heston: a reverse (adjoint) sweep.  The forward pass stores the state
//   (log-spot, variance) every SEG points (16 by default, 32 also built); the
//   backward pass walks the
//   segments from last to first, recomputes each one from its checkpoint
//   (normals regenerated from the Philox counters, kept in registers with the
//   variance), then runs the adjoint recursion of the Euler step back through
//   it, propagating d/dv through sqrt and the floor and accumulating
//   d/d{kappa, theta, sigma_v, rho}.  The spot seeds of each point (running
//   sums: every fed point; running extremes: the point that set the extreme)
//   enter as the sweep passes them.
// * GBM and correlated GBM: d x_{p+1} / d x_p = 1, so the adjoint sweep is a
//   running sum and collapses into the forward pass: it carries d x/d sigma and
//   d x/d mu per asset (two FMAs per step) and each accumulator's derivative.
#include <type_traits>

#include "kt_path_ops.h"

namespace kt {

namespace {

constexpr float kLn2f = 0.693147180559945309f;
constexpr float kLog2ef = 1.44269504088896341f;

template <class AA>
__device__ __forceinline__ int acc_index(const AA& A, uint32_t slot) {
  int q = 0;
#pragma unroll
  for (int i = 1; i < kAdjMaxAcc; ++i)
    if ((uint32_t)A.acc_slot[i] == slot) q = i;
  return q;
}

// Dual-number pass over the ops of point k: every register carries its value
// and its derivative with respect to one leaf (leaf < K: the spot of asset
// `leaf` at this point; leaf = K + q: accumulator q).  Accumulator registers
// keep their forward values (their update ops only set the leaf's tangent),
// every other register is recomputed.  Returns d(pv)/d(leaf) through EMIT.
template <int NS, int K, class AA, class RF>
__device__ float dual_point(const AA& A, int k, int begin, int count, const float (&S)[K], const RF& R,
                            int leaf) {
  const auto& P = A.p;
  float rv[NS], rt[NS];
#pragma unroll
  for (int i = 0; i < NS; ++i) {
    rv[i] = R(0, (uint32_t)i);
    rt[i] = 0.0f;
  }
  float dpv = 0.0f;
  const float df = (float)ld_f64(P.df, (size_t)k);
  for (int i = 0; i < count; ++i) {
    const uint64_t word = ld_u64(P.ops, (size_t)begin + i);
    const uint32_t wx = (uint32_t)word;
    const uint32_t code = wx & 0xffu;
    const uint32_t d = (wx >> 8) & (NS - 1), a = (wx >> 16) & (NS - 1), b = (wx >> 24) & (NS - 1);
    const float imm = __uint_as_float((uint32_t)(word >> 32));
    const float va = rv[a], vb = rv[b], ta = rt[a], tb = rt[b];
    float v = 0.0f, t = 0.0f;
    bool wr = true;
    switch (code) {
      case OP_CONST: v = imm; break;
      case OP_SPOT: {
        const uint32_t asset = (wx >> 16) & 0xffu;
        const uint32_t ai = asset < (uint32_t)K ? asset : 0u;
        v = S[0];
#pragma unroll
        for (int c = 1; c < K; ++c)
          if (ai == (uint32_t)c) v = S[c];
        t = (uint32_t)leaf == ai ? 1.0f : 0.0f;
        break;
      }
      case OP_NEG: v = -va; t = -ta; break;
      case OP_ADD: v = va + vb; t = ta + tb; break;
      case OP_SUB: v = va - vb; t = ta - tb; break;
      case OP_MUL: v = va * vb; t = fmaf(ta, vb, va * tb); break;
      case OP_DIV: v = va / vb; t = (ta - v * tb) / vb; break;
      case OP_MAX: v = va >= vb ? va : vb; t = va >= vb ? ta : tb; break;
      case OP_MIN: v = va <= vb ? va : vb; t = va <= vb ? ta : tb; break;
      case OP_ADDI: v = va + imm; t = ta; break;
      case OP_SUBI: v = va - imm; t = ta; break;
      case OP_RSUBI: v = imm - va; t = -ta; break;
      case OP_MULI: v = va * imm; t = ta * imm; break;
      case OP_DIVI: v = va / imm; t = ta / imm; break;
      case OP_RDIVI: v = imm / va; t = -v * ta / va; break;
      case OP_MAXI: v = va >= imm ? va : imm; t = va >= imm ? ta : 0.0f; break;
      case OP_MINI: v = va <= imm ? va : imm; t = va <= imm ? ta : 0.0f; break;
      case OP_AVG:
      case OP_RMAX:
      case OP_RMIN: {
        const int q = acc_index(A, d);
        if (code == OP_AVG) rt[(d + 1) & (NS - 1)] = 0.0f;  // compensation term: no derivative
        v = rv[d];
        t = leaf == K + q ? 1.0f : 0.0f;
        break;
      }
      case OP_EMIT:
        dpv = fmaf(imm * fmaf(ta, vb, va * tb), df, dpv);
        wr = false;
        break;
      default: wr = false; break;
    }
    if (wr) {
      rv[d] = v;
      rt[d] = t;
    }
  }
  return dpv;
}

// Per-block reduction of NOUT per-path samples into partials[blk][NOUT].
template <int NOUT>
__device__ __forceinline__ void block_stats_out(const double (&vals)[NOUT], bool live, KStats* partials,
                                                uint32_t blk) {
  __shared__ KStats sh[kBlock / 64][NOUT];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int j = 0; j < NOUT; ++j) {
    const KStats s = wave_stats(vals[j], live);
    if (lane == 0) sh[wave][j] = s;
  }
  __syncthreads();
  if (threadIdx.x < NOUT) {
    KStats t = sh[0][threadIdx.x];
#pragma unroll
    for (int wv = 1; wv < kBlock / 64; ++wv) t = stats_combine(t, sh[wv][threadIdx.x]);
    partials[(size_t)blk * NOUT + threadIdx.x] = t;
  }
  __syncthreads();  // sh is reused by the block's next work item
}

// Forward part shared by both kernels: one grid point's program ops, with
// the point at which each running extreme was last set.
template <int NS, int K, class AA, class RF>
__device__ __forceinline__ void run_point_ops(const AA& A, int k, const RF& R, double (&pv)[1][1],
                                              const float (&x)[1][K], bool live, uint32_t fm,
                                              int (&argstep)[kAdjMaxAcc]) {
  const uint64_t st = ld_u64(A.p.steps, (size_t)k);
  const int begin = (int)(uint32_t)st, count = (int)(uint32_t)(st >> 32);
  if (count == 0) return;
  float prev[kAdjMaxAcc];
#pragma unroll
  for (int q = 0; q < kAdjMaxAcc; ++q) prev[q] = (fm >> q) & 1u ? R(0, A.acc_slot[q] & (NS - 1)) : 0.0f;
  run_ops_rf<NS, 1, K, 1>(A.p, k, begin, count, R, pv, x, live);
#pragma unroll
  for (int q = 0; q < kAdjMaxAcc; ++q)
    if (((fm >> q) & 1u) && A.acc_kind[q] != OP_AVG && R(0, A.acc_slot[q] & (NS - 1)) != prev[q]) argstep[q] = k;
}

// A point of a G_AVG_AT group whose mask bit is set: its op list is the
// spot of one asset and that spot's Kahan update of a running sum; run_ops'
// SPOT and AVG expressions on the same slots without the dispatch, rounded
// the same way (the spot's own slot is not written: no op reads it after
// that point's list).  An AVG
// accumulator never moves an argstep, so run_point_ops' bookkeeping has
// nothing to do here.
template <int NS, int K, class AA, class RF>
__device__ __forceinline__ void fix_point(const AA& A, const RF& R, uint32_t slot, uint32_t asset,
                                          const float (&x)[1][K]) {
  // the interpreter rounds the spot to a register slot before the update
  // reads it: no fused multiply-subtract here either
#pragma clang fp contract(off)
  const uint32_t ai = asset < (uint32_t)K ? asset : 0u;
  const float S = scen_s0<K, 1>(A.p, 0, ai) * exp2_fast(x[0][ai]);
  const uint32_t d = slot & (NS - 1), dc = (d + 1) & (NS - 1);
  const float y = S - R(0, dc);
  const float t = R(0, d) + y;
  R(0, dc) = (t - R(0, d)) - y;
  R(0, d) = t;
}

// The mask, slot and asset of group g when it is a G_AVG_AT group, else a
// zero mask (its points take run_point_ops).
struct FixGroup {
  uint32_t mask, slot, asset;
};
template <class PA>
__device__ __forceinline__ FixGroup fix_group(const PA& P, int g) {
  if (ld_u32(P.groups, (size_t)g * 4) != (uint32_t)G_AVG_AT) return FixGroup{0u, 0u, 0u};
  return FixGroup{ld_u32(P.groups, (size_t)g * 4 + 3), ld_u32(P.groups, (size_t)g * 4 + 1),
                  ld_u32(P.groups, (size_t)g * 4 + 2)};
}

template <int NS, class PA, class RF>
__device__ __forceinline__ void init_registers(const PA& P, const RF& R) {
#pragma unroll
  for (int i = 0; i < NS; ++i) R(0, (uint32_t)i) = 0.0f;
  for (int i = 0; i < P.n_init; ++i) {
    const uint64_t w = ld_u64(P.init, (size_t)i);
    R(0, ((uint32_t)w >> 8) & (NS - 1)) = __uint_as_float((uint32_t)(w >> 32));
  }
}

// The op register file of the Greeks kernels, in LDS or in VGPRs.  In LDS
// (NS slots per lane, one column per lane: slot i of lane l at
// i * kBlock + l, so a wave's access is conflict-free) an op of the
// interpreter is a load or a store at a wave-uniform slot index; in VGPRs
// the interpreter's writes at such an index merge whole copies of the file.
// The GBM and Heston adjoints interpret ops at every fixing of a running
// average and take the LDS file (measured 4% faster on GBM Asians and
// Europeans together, 2% on Heston Asians); the basket adjoint takes it
// too, for the registers it frees (at 6 waves per SIMD its spills shrink
// from 320 to 196 B per lane: 1% faster than the VGPR file at 5 waves);
// the tangent kernel reads its file on every point and keeps VGPRs.  Each lane
// touches only its own column, so no barrier is needed between work items.
template <bool IN_LDS, int N>
__device__ __forceinline__ auto adj_register_file(float* lds, float (&arr)[1][N]) {
  if constexpr (IN_LDS) return RegLds{lds + threadIdx.x, kBlock};
  else return RegArray<1, N>{arr};
}
#define KT_ADJ_REGISTER_FILE(NS, R, IN_LDS)                \
  __shared__ float R##_lds[(IN_LDS) ? (NS) * kBlock : 1]; \
  float R##_arr[1][(IN_LDS) ? 1 : (NS)];                  \
  const auto R = adj_register_file<(IN_LDS)>(R##_lds, R##_arr)

}  // namespace

// GBM (K = 1) and correlated GBM (K >= 2): block blk of the job (paths
// blk*256 .. blk*256+255).  Outputs: price, d/ds0_a, d/dsigma_a, d/dmu_a.
template <int K, int NS, class AA>
__device__ __forceinline__ void adjoint_gbm_block(const AA& A, uint32_t blk) {
  constexpr int MODEL = K == 1 ? MK_GBM : MK_BASKET;
  constexpr int M = K;
  constexpr int NP = 3 * K, NOUT = NP + 1;
  const auto& P = A.p;
  const uint64_t local = (uint64_t)blk * kBlock + threadIdx.x;
  const bool live = local < P.n_paths;
  const uint64_t path = P.path_base + local;
  const uint32_t p_lo = (uint32_t)path, p_hi = (uint32_t)(path >> 32);
  const int n = P.n_grid - 1;
  KT_ADJ_REGISTER_FILE(NS, r, true);
  init_registers<NS>(P, r);
  // ws[a] = sum of sqrt(dt) w_a over the steps so far: d x_a / d sigma_a at
  // time t is (ws[a] - sigma_a t) log2(e) and d x_a / d mu_a is t log2(e),
  // so one FMA per asset and point carries both.
  float x[1][K], ws[K];
#pragma unroll
  for (int a = 0; a < K; ++a) x[0][a] = ws[a] = 0.0f;
  float W[kAdjMaxAcc][3];
  int argstep[kAdjMaxAcc];
#pragma unroll
  for (int q = 0; q < kAdjMaxAcc; ++q) {
    W[q][0] = W[q][1] = W[q][2] = 0.0f;
    argstep[q] = 0;
  }
  double pv[1][1] = {{0.0}};
  // Groups of four points without ops (G_NONE: nothing evaluated, nothing
  // fed) only advance the log-spots and the Brownian sums, with the
  // per-point path's expressions on the same values.  Under GBM every run of
  // them is a loop of its own (the register file and the accumulators'
  // state stay where they are, no copies per group); on a grid of equal
  // steps it takes the step coefficients and sqrt(dt) from VGPR copies of
  // point 1's rows (FP32 instructions with SGPR operands issue at about half
  // rate), otherwise it reads the table.
  auto quiet_group = [&](int g, auto&& coef) {
    float z[4 * M];
    group_normals<M>(g, p_lo, p_hi, P.seed_lo, P.seed_hi, 0u, z);
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const int k = 4 * g + s;
      const float* zk = &z[s * M];
      float sq, ta[K], tb[K], ch[K * (K + 1) / 2];
      coef(k, sq, ta, tb, ch);
#pragma unroll
      for (int a = 0; a < K; ++a) {
        float w = 0.0f;
#pragma unroll
        for (int c = 0; c <= a; ++c) w = fmaf(ch[a * (a + 1) / 2 + c], zk[c], w);
        ws[a] = fmaf(sq, w, ws[a]);
        x[0][a] = fmaf(tb[a], w, x[0][a] + ta[a]);
      }
    }
  };
  auto table_coef = [&](int k, float& sq, float (&ta)[K], float (&tb)[K], float (&ch)[K * (K + 1) / 2]) {
    sq = ld_f32(A.adj_tab, 2 * (size_t)k);
#pragma unroll
    for (int a = 0; a < K; ++a) {
      ta[a] = ld_f32(P.tab, ((size_t)k * K + a) * 2);
      tb[a] = ld_f32(P.tab, ((size_t)k * K + a) * 2 + 1);
#pragma unroll
      for (int c = 0; c <= a; ++c) ch[a * (a + 1) / 2 + c] = K == 1 ? 1.0f : P.chol[a * K + c];
    }
  };
  // the last group may be partial: it always takes the per-point path
  auto op_free = [&](int g) {
    return g + 1 < P.n_groups && ld_u32(P.groups, (size_t)g * 4) == (uint32_t)G_NONE;
  };
  int g = 0;
  // (GBM only: with the basket's 26 copies the kernel spills; its price
  // kernel gained 2% from them)
  const bool urow = K == 1 && P.tab_uniform;
  float usq = 0.0f, ua[K], ub[K], uch[K * (K + 1) / 2];
  if (urow) {
    usq = vgpr_copy(ld_f32(A.adj_tab, 2));
#pragma unroll
    for (int a = 0; a < K; ++a) {
      ua[a] = vgpr_copy(ld_f32(P.tab, ((size_t)K + a) * 2));
      ub[a] = vgpr_copy(ld_f32(P.tab, ((size_t)K + a) * 2 + 1));
      // (one asset: the factor is the literal 1, which the compiler folds away)
#pragma unroll
      for (int c = 0; c <= a; ++c) uch[a * (a + 1) / 2 + c] = K == 1 ? 1.0f : vgpr_copy(P.chol[a * K + c]);
    }
  }
  auto row_coef = [&](int, float& sq, float (&ta)[K], float (&tb)[K], float (&ch)[K * (K + 1) / 2]) {
    sq = usq;
#pragma unroll
    for (int a = 0; a < K; ++a) {
      ta[a] = ua[a];
      tb[a] = ub[a];
    }
#pragma unroll
    for (int i = 0; i < K * (K + 1) / 2; ++i) ch[i] = uch[i];
  };
  // (a loop of their own needs a second instance of the basket's Cholesky
  // step, which costs it occupancy: its op-free groups take the per-point
  // path and skip the ops)
  constexpr bool kRuns = K == 1;
  // group 0 holds point 0's zero row: it reads the table
  if (kRuns && op_free(0)) {
    quiet_group(0, table_coef);
    g = 1;
  }
  for (;; ++g) {
    if (kRuns) {
      const int e = op_free_run_end(P, g);
      if (urow)
        for (; g < e; ++g) quiet_group(g, row_coef);
      else
        for (; g < e; ++g) quiet_group(g, table_coef);
    }
    if (g >= P.n_groups) break;
    const bool quiet = !kRuns && op_free(g);
    // (one asset only: the basket's extra code path costs it spills, and its
    // fixings take the interpreter)
    const FixGroup fg = K == 1 ? fix_group(P, g) : FixGroup{0u, 0u, 0u};
    float z[4 * M];
#pragma unroll
    for (int bk = 0; bk < M; ++bk) {
      const uint4 q = philox4x32_10((uint32_t)(g * M + bk), p_lo, p_hi, 0u, P.seed_lo, P.seed_hi);
      z[4 * bk + 0] = normal_from_word(q.x);
      z[4 * bk + 1] = normal_from_word(q.y);
      z[4 * bk + 2] = normal_from_word(q.z);
      z[4 * bk + 3] = normal_from_word(q.w);
    }
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const int k = 4 * g + s;
      if (k <= n) {
        const float* zk = &z[s * M];
        const float sq = ld_f32(A.adj_tab, 2 * (size_t)k);
#pragma unroll
        for (int a = 0; a < K; ++a) {
          float w = 0.0f;
#pragma unroll
          for (int c = 0; c <= a; ++c) w = fmaf(K == 1 ? 1.0f : P.chol[a * K + c], zk[c], w);
          ws[a] = fmaf(sq, w, ws[a]);
          // advance's exact log step on the same w (computed once)
          x[0][a] = fmaf(ld_f32(P.tab, ((size_t)k * K + a) * 2 + 1), w, x[0][a] + ld_f32(P.tab, ((size_t)k * K + a) * 2));
        }
        if (quiet) continue;
        const uint32_t fm = ld_u32(A.feed, (size_t)k);
        int before[kAdjMaxAcc];
#pragma unroll
        for (int q = 0; q < kAdjMaxAcc; ++q) before[q] = argstep[q];
        if (K > 1 || fg.mask == 0u) run_point_ops<NS, K>(A, k, r, pv, x, live, fm, argstep);
        else if ((fg.mask >> s) & 1u) fix_point<NS, K>(A, r, fg.slot, fg.asset, x);
        if (fm) {
#pragma unroll
          for (int q = 0; q < kAdjMaxAcc; ++q) {
            if ((fm >> q) & 1u) {
              const int aq = A.acc_asset[q];
              float xa = x[0][0], wa = ws[0], sga = A.sigma[0], s0a = P.s0[0];
#pragma unroll
              for (int c = 1; c < K; ++c)
                if (aq == c) {
                  xa = x[0][c];
                  wa = ws[c];
                  sga = A.sigma[c];
                  s0a = P.s0[c];
                }
              const float jma = ld_f32(A.adj_tab, 2 * (size_t)k + 1);  // t log2(e)
              const float ja = fmaf(wa, kLog2ef, -sga * jma);
              // S / s0 = 2^x: no division on the per-point path
              const float g = exp2_fast(xa), S = s0a * g;
              const float d0 = S * kLn2f * ja, d1 = S * kLn2f * jma, d2 = g;
              if (A.acc_kind[q] == OP_AVG) {
                W[q][0] += d0;
                W[q][1] += d1;
                W[q][2] += d2;
              } else if (argstep[q] != before[q] || k == 0) {
                W[q][0] = d0;
                W[q][1] = d1;
                W[q][2] = d2;
              }
            }
          }
        }
      }
    }
  }
  // payoff derivatives at the flow date
  const float tl_n = ld_f32(A.adj_tab, 2 * (size_t)n + 1);
  float S[K], js[K];
#pragma unroll
  for (int a = 0; a < K; ++a) js[a] = fmaf(ws[a], kLog2ef, -A.sigma[a] * tl_n);
#pragma unroll
  for (int a = 0; a < K; ++a) S[a] = P.s0[a] * exp2_fast(x[0][a]);
  float gr[NP];
#pragma unroll
  for (int i = 0; i < NP; ++i) gr[i] = 0.0f;
  const uint64_t st = ld_u64(P.steps, (size_t)n);
  const int begin = (int)(uint32_t)st, count = (int)(uint32_t)(st >> 32);
  if (count > 0) {
#pragma unroll
    for (int a = 0; a < K; ++a) {
      const float ds = dual_point<NS, K>(A, n, begin, count, S, r, a);
      gr[a] += ds * S[a] / P.s0[a];
      gr[K + a] += ds * S[a] * kLn2f * js[a];
      gr[2 * K + a] += ds * S[a] * kLn2f * tl_n;
    }
    for (int q = 0; q < A.n_acc; ++q) {
      const float ab = dual_point<NS, K>(A, n, begin, count, S, r, K + q);
      float w0 = W[0][0], w1 = W[0][1], w2 = W[0][2];
#pragma unroll
      for (int c = 1; c < kAdjMaxAcc; ++c)
        if (q == c) {
          w0 = W[c][0];
          w1 = W[c][1];
          w2 = W[c][2];
        }
      const int aq = A.acc_asset[q];
#pragma unroll
      for (int a = 0; a < K; ++a)
        if (aq == a) {
          gr[a] += ab * w2;
          gr[K + a] += ab * w0;
          gr[2 * K + a] += ab * w1;
        }
    }
  }
  double vals[NOUT];
  vals[0] = pv[0][0];
#pragma unroll
  for (int i = 0; i < NP; ++i) vals[1 + i] = (double)gr[i];
  block_stats_out<NOUT>(vals, live, P.partials, blk);
}

template <int K, int NS>
__global__ __launch_bounds__(kBlock) void adjoint_gbm_kernel(const AdjArgs A) {
  adjoint_gbm_block<K, NS>(A, blockIdx.x);
}

// One Heston Euler step of the variance only (the same expression as
// advance<MK_HESTON>, so the recomputed variance is bit-identical).
template <class PA>
__device__ __forceinline__ float heston_v_step(const PA& P, int k, float z1, float z2, float v) {
  const float t3 = ld_f32(P.tab, (size_t)kHestonStride * k + 3);
  const float t4 = ld_f32(P.tab, (size_t)kHestonStride * k + 4);
  const float t5 = ld_f32(P.tab, (size_t)kHestonStride * k + 5);
  const float vp = fmaxf(v, 0.0f);
  const float sv = __builtin_amdgcn_sqrtf(vp);
  return heston_v_next(v, vp, sv, z1, z2, t3, t4, t5, P.rho, P.rho_perp, heston_floor(P));
}

// Heston in forward (tangent) mode: one pass carrying d x / d theta and
// d v / d theta for theta in {v0, kappa, theta, sigma_v, rho} through the same
// Euler step, sqrt and floor as the reverse sweep (its transpose), plus each
// accumulator's derivatives at its fed points.  With five variance
// parameters one forward pass of tangents costs less than the reverse
// sweep's second generation of every normal; the estimator is the same
// pathwise derivative, to rounding.  Outputs as the reverse kernel.
template <int NS, bool FEEDS, class AA>
__device__ __forceinline__ void tangent_heston_block(const AA& A, uint32_t blk) {
  constexpr int NOUT = 7, NT = 5;  // tangents: v0, kappa, theta, sigma_v, rho
  const auto& P = A.p;
  const uint64_t local = (uint64_t)blk * kBlock + threadIdx.x;
  const bool live = local < P.n_paths;
  const uint64_t path = P.path_base + local;
  const uint32_t p_lo = (uint32_t)path, p_hi = (uint32_t)(path >> 32);
  const int n = P.n_grid - 1;
  KT_ADJ_REGISTER_FILE(NS, r, false);
  init_registers<NS>(P, r);
  float x[1][1] = {{0.0f}};
  float v = P.v0;
  int argstep[kAdjMaxAcc];
#pragma unroll
  for (int q = 0; q < kAdjMaxAcc; ++q) argstep[q] = 0;
  double pv[1][1] = {{0.0}};
  const bool ft = P.heston_full_truncation != 0;
  const float theta = A.theta, rho = P.rho, rho_perp = P.rho_perp;
  const float rho_ratio = rho_perp > 0.0f ? rho / rho_perp : 0.0f;
  const float s0 = P.s0[0];
  float dx[NT], dv[NT];
#pragma unroll
  for (int j = 0; j < NT; ++j) dx[j] = dv[j] = 0.0f;
  dv[0] = P.v0 > 0.0f ? 1.0f : 0.0f;
  float W[FEEDS ? kAdjMaxAcc : 1][NT + 1];  // per accumulator: d/ds0, then the tangents
#pragma unroll
  for (int q = 0; q < (FEEDS ? kAdjMaxAcc : 1); ++q)
#pragma unroll
    for (int j = 0; j <= NT; ++j) W[q][j] = 0.0f;
  // Step coefficients of a grid of equal steps (rows 1 .. n shared): VGPR
  // copies for the leading op-free groups (FP32 instructions with SGPR
  // operands issue at about half rate).  Filled only when P.tab_uniform.
  StepRow<MK_HESTON, 1, 1> urow;
  float u_rev[7], u_theta = 0.0f, u_rr = 0.0f;
#pragma unroll
  for (int q = 0; q < 7; ++q) u_rev[q] = 0.0f;
  // one Euler step at point k >= 1 with its tangents; U: coefficients from
  // the VGPR copies (the same values, so the same results)
  auto step = [&](int k, const float* zk, auto use_row) {
    constexpr bool U = decltype(use_row)::value;
    const float z1 = zk[0], z2 = zk[1];
    if (k >= 1) {
      const float t1 = U ? u_rev[0] : ld_f32(A.rev, 8 * (size_t)k + 0);
      const float t2 = U ? u_rev[1] : ld_f32(A.rev, 8 * (size_t)k + 1);
      const float t3 = U ? u_rev[2] : ld_f32(A.rev, 8 * (size_t)k + 2);
      const float t5 = U ? u_rev[3] : ld_f32(A.rev, 8 * (size_t)k + 3);
      const float dt = U ? u_rev[4] : ld_f32(A.rev, 8 * (size_t)k + 4);
      const float sq = U ? u_rev[5] : ld_f32(A.rev, 8 * (size_t)k + 5);
      const float kdt = U ? u_rev[6] : ld_f32(A.rev, 8 * (size_t)k + 6);
      const float th = U ? u_theta : theta, rr = U ? u_rr : rho_ratio;
      const float vp = fmaxf(v, 0.0f);  // = v under the floor scheme
      const float sv = __builtin_amdgcn_sqrtf(vp);
      const float w = U ? fmaf(urow.rho, z1, urow.rho_perp * z2) : fmaf(rho, z1, rho_perp * z2);
      const bool pass = !ft || v > 0.0f;  // d vp / d v
      const float hs = sv > 0.0f ? 0.5f * __builtin_amdgcn_rsqf(vp) : 0.0f;  // d sqrt(vp) / d vp
      const float cx = t2 * z1, cv = t5 * w;
      float dvn[NT];
#pragma unroll
      for (int j = 0; j < NT; ++j) {
        const float dvp = pass ? dv[j] : 0.0f;
        const float dsv = hs * dvp;
        dx[j] = fmaf(cx, dsv, fmaf(t1, dvp, dx[j]));
        dvn[j] = fmaf(cv, dsv, fmaf(t3, dvp, dv[j]));
      }
      dvn[1] = fmaf(dt, th - vp, dvn[1]);                // kappa
      dvn[2] += kdt;                                     // theta
      dvn[3] = fmaf(sq * sv, w, dvn[3]);                 // sigma_v
      dvn[4] = fmaf(t5 * sv, fmaf(-rr, z2, z1), dvn[4]);  // rho
      if constexpr (U) advance_row<MK_HESTON, 1, 1>(urow, zk, x, v);
      else advance<MK_HESTON, 1, 1>(P, k, zk, x, v);
      const bool keep = ft || v > 0.0f;  // the floor cuts the derivative
#pragma unroll
      for (int j = 0; j < NT; ++j) dv[j] = keep ? dvn[j] : 0.0f;
    }
  };
  // leading groups whose points run no ops and feed nothing (P.n_uniform);
  // on a grid of equal steps groups 1 .. n_quiet - 1 take the VGPR rows
  // (group 0 holds point 0's zero row)
  const int n_quiet = min(P.n_uniform, P.n_groups);
  const std::integral_constant<bool, true> kRow;
  const std::integral_constant<bool, false> kTable;
  int g_row = 0;
  if (P.tab_uniform && n_quiet > 1) {
    {
      float z[8];
      group_normals<2>(0, p_lo, p_hi, P.seed_lo, P.seed_hi, 0u, z);
#pragma unroll
      for (int s = 0; s < 4; ++s) step(s, &z[2 * s], kTable);
    }
    urow = vgpr_row<MK_HESTON, 1, 1>(P, 1);
#pragma unroll
    for (int q = 0; q < 7; ++q) u_rev[q] = vgpr_copy(ld_f32(A.rev, 8 + q));
    u_theta = vgpr_copy(theta);
    u_rr = vgpr_copy(rho_ratio);
    for (int g = 1; g < n_quiet; ++g) {
      float z[8];
      group_normals<2>(g, p_lo, p_hi, P.seed_lo, P.seed_hi, 0u, z);
#pragma unroll
      for (int s = 0; s < 4; ++s) step(4 * g + s, &z[2 * s], kRow);
    }
    g_row = n_quiet;
  }
  for (int g = g_row; g < n_quiet; ++g) {
    float z[8];
    group_normals<2>(g, p_lo, p_hi, P.seed_lo, P.seed_hi, 0u, z);
#pragma unroll
    for (int s = 0; s < 4; ++s) step(4 * g + s, &z[2 * s], kTable);
  }
  for (int g = n_quiet; g < P.n_groups; ++g) {
    float z[8];
    group_normals<2>(g, p_lo, p_hi, P.seed_lo, P.seed_hi, 0u, z);
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const int k = 4 * g + s;
      if (k <= n) {
        step(k, &z[2 * s], kTable);
        if constexpr (FEEDS) {
          const uint32_t fm = ld_u32(A.feed, (size_t)k);
          int before[kAdjMaxAcc];
#pragma unroll
          for (int q = 0; q < kAdjMaxAcc; ++q) before[q] = argstep[q];
          run_point_ops<NS, 1>(A, k, r, pv, x, live, fm, argstep);
          if (fm) {
            const float e = exp2_fast(x[0][0]), S = s0 * e, sl = S * kLn2f;
#pragma unroll
            for (int q = 0; q < kAdjMaxAcc; ++q) {
              if ((fm >> q) & 1u) {
                if (A.acc_kind[q] == OP_AVG) {
                  W[q][0] += e;
#pragma unroll
                  for (int j = 0; j < NT; ++j) W[q][1 + j] = fmaf(sl, dx[j], W[q][1 + j]);
                } else if (argstep[q] != before[q] || k == 0) {
                  W[q][0] = e;
#pragma unroll
                  for (int j = 0; j < NT; ++j) W[q][1 + j] = sl * dx[j];
                }
              }
            }
          }
        } else {
          run_point_ops<NS, 1>(A, k, r, pv, x, live, 0u, argstep);
        }
      }
    }
  }
  const float S_n = s0 * exp2_fast(x[0][0]);
  float g[NT + 1];
#pragma unroll
  for (int j = 0; j <= NT; ++j) g[j] = 0.0f;
  const uint64_t st = ld_u64(P.steps, (size_t)n);
  const int begin = (int)(uint32_t)st, count = (int)(uint32_t)(st >> 32);
  if (count > 0) {
    const float S[1] = {S_n};
    const float ds_n = dual_point<NS, 1>(A, n, begin, count, S, r, 0);
    g[0] = ds_n * S_n / s0;
#pragma unroll
    for (int j = 0; j < NT; ++j) g[1 + j] = ds_n * S_n * kLn2f * dx[j];
    if constexpr (FEEDS) {
#pragma unroll
      for (int q = 0; q < kAdjMaxAcc; ++q) {
        if (q < A.n_acc) {
          const float ab = dual_point<NS, 1>(A, n, begin, count, S, r, 1 + q);
#pragma unroll
          for (int j = 0; j <= NT; ++j) g[j] = fmaf(ab, W[q][j], g[j]);
        }
      }
    }
  }
  double vals[NOUT] = {pv[0][0], (double)g[0], (double)g[1], (double)g[2], (double)g[3], (double)g[4], (double)g[5]};
  block_stats_out<NOUT>(vals, live, P.partials, blk);
}

template <int NS, bool FEEDS>
__global__ __launch_bounds__(kBlock) void tangent_heston_kernel(const AdjArgs A) {
  tangent_heston_block<NS, FEEDS>(A, blockIdx.x);
}

// Heston: forward with a checkpoint every SEG points, then the adjoint sweep
// with per-segment recompute.  Outputs: price, d/d{s0, v0, kappa, theta,
// sigma_v, rho}.  FEEDS: the payoff has running averages / extremes, so the
// sweep needs the log-spot at every point (recomputed with the variance and
// walked back by inverting the log step); without them only the last point
// seeds d/dx, which then stays constant, and the sweep carries the variance
// alone.
// Block blk of the job; this lane's checkpoints are ck[(2c + i) * cs + lane].
template <int NS, int SEG, bool FEEDS, class AA>
__device__ __forceinline__ void adjoint_heston_block(const AA& A, uint32_t blk, float* __restrict__ ck, size_t cs,
                                                     size_t lane) {
  constexpr int NOUT = 7;
  const auto& P = A.p;
  const uint64_t local = (uint64_t)blk * kBlock + threadIdx.x;
  const bool live = local < P.n_paths;
  const uint64_t path = P.path_base + local;
  const uint32_t p_lo = (uint32_t)path, p_hi = (uint32_t)(path >> 32);
  const int n = P.n_grid - 1;
  KT_ADJ_REGISTER_FILE(NS, r, true);
  init_registers<NS>(P, r);
  float x[1][1] = {{0.0f}};
  float v = P.v0;
  int argstep[kAdjMaxAcc];
#pragma unroll
  for (int q = 0; q < kAdjMaxAcc; ++q) argstep[q] = 0;
  double pv[1][1] = {{0.0}};
  auto checkpoint = [&](int k) {
    if ((k & (SEG - 1)) == 0 && live) {
      const size_t c = (size_t)(k / SEG);
      if constexpr (FEEDS) ck[(2 * c) * cs + lane] = x[0][0];
      ck[(2 * c + 1) * cs + lane] = v;
    }
  };
  // leading groups without ops (P.n_uniform): advance and checkpoint only.
  // On a grid of equal steps groups 1 .. n_quiet - 1 take the step row from
  // VGPR copies of row 1 (group 0 holds point 0's zero row and reads the
  // table): the same expressions on the same values as advance().
  const int n_quiet = min(P.n_uniform, P.n_groups);
  int g_row = 0;
  if (P.tab_uniform && n_quiet > 1) {
    {
      float z[8];
      group_normals<2>(0, p_lo, p_hi, P.seed_lo, P.seed_hi, 0u, z);
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        checkpoint(s);
        advance<MK_HESTON, 1, 1>(P, s, &z[2 * s], x, v);
      }
    }
    const StepRow<MK_HESTON, 1, 1> R = vgpr_row<MK_HESTON, 1, 1>(P, 1);
    for (int g = 1; g < n_quiet; ++g) {
      float z[8];
      group_normals<2>(g, p_lo, p_hi, P.seed_lo, P.seed_hi, 0u, z);
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        checkpoint(4 * g + s);
        advance_row<MK_HESTON, 1, 1>(R, &z[2 * s], x, v);
      }
    }
    g_row = n_quiet;
  }
  for (int g = g_row; g < n_quiet; ++g) {
    float z[8];
    group_normals<2>(g, p_lo, p_hi, P.seed_lo, P.seed_hi, 0u, z);
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      checkpoint(4 * g + s);
      advance<MK_HESTON, 1, 1>(P, 4 * g + s, &z[2 * s], x, v);
    }
  }
  for (int g = n_quiet; g < P.n_groups; ++g) {
    // later runs of groups without ops (between the fixings of a running
    // average): on a grid of equal steps from VGPR copies of row 1 as well
    if (P.tab_uniform && g > 0) {
      const int e = op_free_run_end(P, g);
      if (e > g) {
        const StepRow<MK_HESTON, 1, 1> R = vgpr_row<MK_HESTON, 1, 1>(P, 1);
        for (; g < e; ++g) {
          float z[8];
          group_normals<2>(g, p_lo, p_hi, P.seed_lo, P.seed_hi, 0u, z);
#pragma unroll
          for (int s = 0; s < 4; ++s) {
            checkpoint(4 * g + s);
            advance_row<MK_HESTON, 1, 1>(R, &z[2 * s], x, v);
          }
        }
      }
    }
    float z[8];
    group_normals<2>(g, p_lo, p_hi, P.seed_lo, P.seed_hi, 0u, z);
    // a later group without ops (full: only the last one can be partial)
    if (g + 1 < P.n_groups && ld_u32(P.groups, (size_t)g * 4) == (uint32_t)G_NONE) {
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        checkpoint(4 * g + s);
        advance<MK_HESTON, 1, 1>(P, 4 * g + s, &z[2 * s], x, v);
      }
      continue;
    }
    const FixGroup fg = fix_group(P, g);
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const int k = 4 * g + s;
      if (k <= n) {
        checkpoint(k);
        advance<MK_HESTON, 1, 1>(P, k, &z[2 * s], x, v);
        if (fg.mask == 0u)
          run_point_ops<NS, 1>(A, k, r, pv, x, live, FEEDS ? ld_u32(A.feed, (size_t)k) : 0u, argstep);
        else if ((fg.mask >> s) & 1u)
          fix_point<NS, 1>(A, r, fg.slot, fg.asset, x);
      }
    }
  }
  // d(pv)/d(spot at n) and d(pv)/d(accumulator q)
  const float s0 = P.s0[0];
  const float S_n = s0 * exp2_fast(x[0][0]);
  float ds_n = 0.0f;
  float abar[kAdjMaxAcc];
#pragma unroll
  for (int q = 0; q < kAdjMaxAcc; ++q) abar[q] = 0.0f;
  {
    const float S[1] = {S_n};
    const uint64_t st = ld_u64(P.steps, (size_t)n);
    const int begin = (int)(uint32_t)st, count = (int)(uint32_t)(st >> 32);
    if (count > 0) {
      ds_n = dual_point<NS, 1>(A, n, begin, count, S, r, 0);
      if constexpr (FEEDS) {
#pragma unroll
        for (int q = 0; q < kAdjMaxAcc; ++q)
          if (q < A.n_acc) abar[q] = dual_point<NS, 1>(A, n, begin, count, S, r, 1 + q);
      }
    }
  }
  const bool ft = P.heston_full_truncation != 0;
  const float theta = A.theta, rho = P.rho, rho_perp = P.rho_perp;
  const float rho_ratio = rho_perp > 0.0f ? rho / rho_perp : 0.0f;
  // without feeds the only spot seed is at n: d/dx is constant along the path
  float xb = FEEDS ? 0.0f : ds_n * S_n * kLn2f;
  float g_s0 = FEEDS ? 0.0f : ds_n * S_n / s0;
  float vbar = 0.0f, g_k = 0.0f, g_th = 0.0f, g_sv = 0.0f, g_rho = 0.0f;
  const int n_seg = n / SEG + 1;
  // checkpoint of the segment to recompute next, loaded one segment ahead so
  // that its latency hides behind the current segment's work
  float x_ck = FEEDS && live ? ck[(2 * (size_t)(n_seg - 1)) * cs + lane] : 0.0f;
  float v_ck = live ? ck[(2 * (size_t)(n_seg - 1) + 1) * cs + lane] : 0.0f;
  // The recompute and the sweep, in two instances: U (a grid of equal steps
  // and no running accumulators) takes the step's table row and the reverse
  // row, shared by points 1 .. n, from VGPR copies (FP32 instructions with
  // SGPR operands issue at about half rate: 5% on config 3), the other reads
  // them per point.  Both evaluate the same expressions on the same values.
  //
  // On the equal-step instance the per-point constants are folded out of the
  // loop: xb t1 and xb t2 / 2 (xb is constant without feeds), t5 / 2 and
  // dt theta are VGPR copies; the theta, sigma_v and rho sums carry
  // sum vnb, sum vnb sqrt(v) w and sum vnb sqrt(v) (z1 - rho z2 / rho'), and
  // are multiplied once at the end by the row's kappa dt, sqrt(dt) and t5.
  auto sweep = [&](auto urow) {
    constexpr bool U = decltype(urow)::value && !FEEDS;
    // with running accumulators on a grid of equal steps: the general
    // expressions, every point's rows read from row 1 (loads the compiler
    // hoists out of the loops)
    constexpr bool UF = decltype(urow)::value && FEEDS;
    float u_e0 = 0.0f, u_d1 = 0.0f, u_d2 = 0.0f, u_rho = 0.0f, u_rhop = 0.0f, u_lo = 0.0f;
    float u_c1 = 0.0f, u_c2h = 0.0f, u_t3 = 0.0f, u_t5h = 0.0f, u_dt = 0.0f, u_dtth = 0.0f, u_rr = 0.0f;
    float s_vnb = 0.0f, s_sv = 0.0f, s_rho = 0.0f;
    if constexpr (U) {
      u_c1 = xb * ld_f32(A.rev, 8 + 0);
      u_c2h = 0.5f * xb * ld_f32(A.rev, 8 + 1);
      u_t3 = vgpr_copy(ld_f32(A.rev, 8 + 2));
      u_t5h = vgpr_copy(0.5f * ld_f32(A.rev, 8 + 3));
      u_dt = vgpr_copy(ld_f32(A.rev, 8 + 4));
      u_dtth = vgpr_copy(ld_f32(A.rev, 8 + 4) * theta);
      u_rr = vgpr_copy(rho_ratio);
      u_e0 = vgpr_copy(ld_f32(P.tab, kHestonStride + 3));
      u_d1 = vgpr_copy(ld_f32(P.tab, kHestonStride + 4));
      u_d2 = vgpr_copy(ld_f32(P.tab, kHestonStride + 5));
      u_rho = vgpr_copy(P.rho);
      u_rhop = vgpr_copy(P.rho_perp);
      u_lo = vgpr_copy(heston_floor(P));
    }
    for (int c = n_seg - 1; c >= 0; --c) {
      // recompute segment c (points SEG*c .. SEG*c + SEG - 1) from its checkpoint
      float xs[1][1] = {{x_ck}};
      float vs = v_ck;
      if (c > 0) {
        if constexpr (FEEDS) x_ck = live ? ck[(2 * (size_t)(c - 1)) * cs + lane] : 0.0f;
        v_ck = live ? ck[(2 * (size_t)(c - 1) + 1) * cs + lane] : 0.0f;
      }
      // per point: sqrt(max(v, 0)) before the step (the sweep needs it and its
      // square, not v), and the two normals
      float svs[SEG], z1s[SEG], z2s[SEG];
#pragma unroll
      for (int gq = 0; gq < SEG / 4; ++gq) {
        const int g = (SEG / 4) * c + gq;
        float z[8];
#pragma unroll
        for (int bk = 0; bk < 2; ++bk) {
          const uint4 q = philox4x32_10((uint32_t)(2 * g + bk), p_lo, p_hi, 0u, P.seed_lo, P.seed_hi);
          z[4 * bk + 0] = normal_from_word(q.x);
          z[4 * bk + 1] = normal_from_word(q.y);
          z[4 * bk + 2] = normal_from_word(q.z);
          z[4 * bk + 3] = normal_from_word(q.w);
        }
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          const int i = 4 * gq + s;
          svs[i] = __builtin_amdgcn_sqrtf(fmaxf(vs, 0.0f));  // the step's own sqrt (same expression)
          z1s[i] = z[2 * s];
          z2s[i] = z[2 * s + 1];
          if (4 * g + s <= n) {
            if constexpr (UF) {
              // point 0's zero row leaves the state as it is
              if (4 * g + s == 0) continue;
              advance<MK_HESTON, 1, 1>(P, 1, &z[2 * s], xs, vs);
            } else if constexpr (FEEDS) advance<MK_HESTON, 1, 1>(P, 4 * g + s, &z[2 * s], xs, vs);
            else if constexpr (U) {
              // point 0 has the zero row (no step); the others share row 1
              if (4 * g + s == 0) continue;
              const float vp0 = fmaxf(vs, 0.0f);
              vs = heston_v_next(vs, vp0, __builtin_amdgcn_sqrtf(vp0), z[2 * s], z[2 * s + 1], u_e0, u_d1, u_d2, u_rho,
                                 u_rhop, u_lo);
            } else vs = heston_v_step(P, 4 * g + s, z[2 * s], z[2 * s + 1], vs);
          }
        }
      }
      // reverse sweep through the segment; xs, vs hold the state after its last point
      float sv_next = __builtin_amdgcn_sqrtf(fmaxf(vs, 0.0f));  // > 0 exactly when v after the step is
#pragma unroll
      for (int i = SEG - 1; i >= 0; --i) {
        const int p = SEG * c + i;
        if (p <= n) {
          if constexpr (FEEDS) {
            // spot seeds only where the payoff reads the spot: the flow date
            // and the fed points (a uniform branch)
            const uint32_t fm = ld_u32(A.feed, (size_t)p);
            if (fm != 0u || p == n) {
              float seed = p == n ? ds_n : 0.0f;
#pragma unroll
              for (int q = 0; q < kAdjMaxAcc; ++q)
                if ((fm >> q) & 1u) seed += (A.acc_kind[q] == OP_AVG || argstep[q] == p) ? abar[q] : 0.0f;
              const float g = exp2_fast(xs[0][0]), S = s0 * g;  // S / s0 = 2^x
              xb = fmaf(seed * S, kLn2f, xb);
              g_s0 = fmaf(seed, g, g_s0);
            }
          }
          if (p >= 1 && U) {
            const float sv = svs[i], z1 = z1s[i], z2 = z2s[i];
            const float vp = sv * sv;  // max(v, 0) to rounding
            const float w = fmaf(u_rho, z1, u_rhop * z2);
            const float vnb = (ft || sv_next > 0.0f) ? vbar : 0.0f;
            const float a = vnb * sv;
            s_vnb += vnb;
            g_k = fmaf(vnb, fmaf(-u_dt, vp, u_dtth), g_k);
            s_sv = fmaf(a, w, s_sv);
            s_rho = fmaf(a, fmaf(-u_rr, z2, z1), s_rho);
            // half of d/d sqrt(v), times 1 / sqrt(v) (hardware reciprocal):
            // d sqrt(v) / dv = 0.5 / sqrt(v)
            const float svbh = fmaf(u_c2h, z1, (vnb * u_t5h) * w);
            const float rs = sv > 0.0f ? __builtin_amdgcn_rcpf(sv) : 0.0f;
            const float vpb = fmaf(svbh, rs, fmaf(vnb, u_t3, u_c1));
            vbar = vnb + ((!ft || sv > 0.0f) ? vpb : 0.0f);
            sv_next = sv;
          } else if (p >= 1) {
            // one 32-byte row per point: a single scalar load
            const size_t pr = UF ? 1 : (size_t)p;
            const float t1 = ld_f32(A.rev, 8 * pr + 0);
            const float t2 = ld_f32(A.rev, 8 * pr + 1);
            const float t3 = ld_f32(A.rev, 8 * pr + 2);
            const float t5 = ld_f32(A.rev, 8 * pr + 3);
            const float dt = ld_f32(A.rev, 8 * pr + 4);
            const float sq = ld_f32(A.rev, 8 * pr + 5);
            const float kdt = ld_f32(A.rev, 8 * pr + 6);
            const float sv = svs[i], z1 = z1s[i], z2 = z2s[i];
            const float vp = sv * sv;  // max(v, 0) to rounding
            const float w = fmaf(rho, z1, rho_perp * z2);
            // v_new = floor(v + e0 vp + d1 + d2 sv w); x_new = x + c0 + c1 vp + c2 sv z1
            const float vnb = (ft || sv_next > 0.0f) ? vbar : 0.0f;
            float vpb = fmaf(xb, t1, vnb * t3);
            const float svb = fmaf(xb * t2, z1, vnb * t5 * w);
            g_k = fmaf(vnb, dt * (theta - vp), g_k);
            g_th = fmaf(vnb, kdt, g_th);
            g_sv = fmaf(vnb, sq * sv * w, g_sv);
            g_rho = fmaf(vnb * t5 * sv, fmaf(-rho_ratio, z2, z1), g_rho);
            // d sqrt(v)/dv = 0.5 / sqrt(v) through the hardware reciprocal (one
            // instruction instead of an IEEE division)
            vpb = sv > 0.0f ? fmaf(svb, 0.5f * __builtin_amdgcn_rcpf(sv), vpb) : vpb;
            vbar = vnb + ((!ft || sv > 0.0f) ? vpb : 0.0f);
            if constexpr (FEEDS) {
              const float t0 = ld_f32(P.tab, (size_t)kHestonStride * pr + 0);
              xs[0][0] = xs[0][0] - (fmaf(t1, vp, t0) + t2 * sv * z1);
            }
            sv_next = sv;
          }
        }
      }
    }
    if constexpr (U) {
      // every row is row 1 (point 0 has no step)
      g_th = ld_f32(A.rev, 8 + 6) * s_vnb;
      g_sv = ld_f32(A.rev, 8 + 5) * s_sv;
      g_rho = ld_f32(A.rev, 8 + 3) * s_rho;
    }
  };
  // (with running accumulators U would fold a constant xb: the equal-step
  // instance is UF)
  if constexpr (!FEEDS) {
    if (P.tab_uniform) sweep(std::integral_constant<bool, true>{});
    else sweep(std::integral_constant<bool, false>{});
  } else {
    if (P.tab_uniform) sweep(std::integral_constant<bool, true>{});
    else sweep(std::integral_constant<bool, false>{});
  }

  double vals[NOUT] = {pv[0][0], (double)g_s0, (double)vbar, (double)g_k, (double)g_th, (double)g_sv, (double)g_rho};
  block_stats_out<NOUT>(vals, live, P.partials, blk);
}

template <int NS, int SEG, bool FEEDS>
__global__ __launch_bounds__(kBlock) void adjoint_heston_kernel(const AdjArgs A) {
  const uint64_t path = A.p.path_base + (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  adjoint_heston_block<NS, SEG, FEEDS>(A, blockIdx.x, A.ck, A.ck_stride, path);
}

// With running averages the default-interval kernel would take ~180 VGPRs
// (3 waves per SIMD); capped at 128 it spills two values and runs 4 waves
// (measured 12% faster on the book's Heston Asians).
template <>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4))) void
adjoint_heston_kernel<16, 16, true>(const AdjArgs A) {
  const uint64_t path = A.p.path_base + (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  adjoint_heston_block<16, 16, true>(A, blockIdx.x, A.ck, A.ck_stride, path);
}

// ---------------------------------------------------------------------------
// Persistent portfolio launch of the pathwise Greeks: one launch per variant
// (each with the register budget of its own block code) takes (job, block)
// items from a queue and runs the same block function as the per-contract
// launch, so every job's statistics equal its kt_greeks_adjoint result.  The
// Heston checkpoints live in a per-workgroup slot (the workgroup finishes an
// item's forward and backward sweeps before it takes the next item).
// ---------------------------------------------------------------------------
template <int VARIANT>
__device__ __forceinline__ void portfolio_adjoint_body(const AdjPortfolioArgs& Q) {
  __shared__ uint32_t s_first;
  const KT_CONST AdjArgs* jobs = (const KT_CONST AdjArgs*)Q.jobs;
  const KT_CONST uint32_t* items = (const KT_CONST uint32_t*)Q.items;
  float* ck = Q.ck ? Q.ck + (size_t)blockIdx.x * Q.ck_floats : nullptr;
  for (;;) {
    __syncthreads();  // every lane has read the previous s_first
    if (threadIdx.x == 0) s_first = atomicAdd(Q.counter, (unsigned)Q.chunk);
    __syncthreads();
    const uint32_t first = __builtin_amdgcn_readfirstlane(s_first);
    if (first >= Q.n_items) return;
    const uint32_t last = min(first + (uint32_t)Q.chunk, Q.n_items);
    for (uint32_t it = first; it < last; ++it) {
      const uint32_t j = items[2 * (size_t)it], blk = items[2 * (size_t)it + 1];
      const KT_CONST AdjArgs& A = jobs[j];
      if constexpr (VARIANT == APV_GBM) adjoint_gbm_block<1, 16>(A, blk);
      if constexpr (VARIANT == APV_HESTON) adjoint_heston_block<16, kAdjSeg, false>(A, blk, ck, kBlock, threadIdx.x);
      if constexpr (VARIANT == APV_HESTON_FEEDS)
        adjoint_heston_block<16, kAdjSeg, true>(A, blk, ck, kBlock, threadIdx.x);
      if constexpr (VARIANT == APV_BASKET5) adjoint_gbm_block<5, 16>(A, blk);
      if constexpr (VARIANT == APV_HESTON_T) tangent_heston_block<16, false>(A, blk);
      if constexpr (VARIANT == APV_HESTON_FEEDS_T) tangent_heston_block<16, true>(A, blk);
    }
  }
}

// The register budget per variant, measured on the 1,000-contract book: the
// Heston variants need a few registers more than 128, which would cost a
// quarter of their occupancy (3 waves per SIMD instead of 4), so they are
// capped at 128 (4 waves; 5 waves was 3% slower); the Heston variant with
// running accumulators runs faster still at 102 (5 waves, spilling a few
// values around its loops: 6%; 6 waves 9% slower), GBM with its register
// file in LDS at 102 too (5 waves, its spills outside the group loop: 6%
// against the compiler's 116 VGPRs), and the basket, with its register
// file in LDS, at 85 (6 waves: 1% against 5 waves).  WAVES = 5 caps every
// variant at 102 VGPRs (a tile the autotuner may pick).
template <int VARIANT>
constexpr int adjoint_default_waves() {
  return VARIANT == APV_BASKET5                                     ? 6
         : (VARIANT == APV_HESTON_FEEDS || VARIANT == APV_GBM)      ? 5
         : (VARIANT == APV_HESTON || VARIANT == APV_HESTON_FEEDS_T) ? 4
                                                                     : 1;
}

template <int VARIANT, int WAVES>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(
    WAVES > 0 ? WAVES : adjoint_default_waves<VARIANT>()))) void portfolio_adjoint_kernel(const AdjPortfolioArgs Q) {
  portfolio_adjoint_body<VARIANT>(Q);
}

// Block b: output b % kAdjMaxOut of job b / kAdjMaxOut (outputs beyond the
// job's count are skipped), into out[job][kAdjMaxOut].
__global__ __launch_bounds__(256) void portfolio_adjoint_finalize_kernel(const AdjArgs* __restrict__ jobs,
                                                                         const int* __restrict__ n_blocks,
                                                                         const int* __restrict__ n_out,
                                                                         KStats* __restrict__ out) {
  const int job = blockIdx.x / kAdjMaxOut, j = blockIdx.x % kAdjMaxOut;
  if (j >= n_out[job]) return;
  finalize_body(jobs[job].p.partials, n_blocks[job], n_out[job], j, out + (size_t)job * kAdjMaxOut + j);
}

int adjoint_outputs(int model, int k_assets) { return model == MK_HESTON ? 7 : 1 + 3 * k_assets; }

bool adjoint_variant_exists(int model, int k_assets, int n_slots, int* ns_chosen) {
  if (n_slots > 32) return false;
  if (model == MK_BASKET && (k_assets < 2 || k_assets > kMaxAssets)) return false;
  *ns_chosen = (n_slots <= 16 && (model != MK_BASKET || k_assets <= 5)) ? 16 : 32;
  return true;
}

int portfolio_adjoint_variant(int model, int k_assets, int n_slots, int n_acc, bool tangent) {
  if (n_slots > 16) return -1;
  if (model == MK_GBM) return APV_GBM;
  if (model == MK_HESTON && tangent) return n_acc > 0 ? APV_HESTON_FEEDS_T : APV_HESTON_T;
  if (model == MK_HESTON) return n_acc > 0 ? APV_HESTON_FEEDS : APV_HESTON;
  if (model == MK_BASKET && k_assets == 5) return APV_BASKET5;
  return -1;
}

namespace {

template <int VARIANT>
int adjoint_occupancy_of(int waves) {
  int n = 0;
  const hipError_t e =
      waves == 5 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, portfolio_adjoint_kernel<VARIANT, 5>, kBlock, 0)
                 : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, portfolio_adjoint_kernel<VARIANT, 0>, kBlock, 0);
  if (e != hipSuccess) return 1;
  return n > 0 ? n : 1;
}

template <int VARIANT>
void launch_adjoint_variant(const AdjPortfolioArgs& q, int grid, hipStream_t s, int waves) {
  if (waves == 5) hipLaunchKernelGGL((portfolio_adjoint_kernel<VARIANT, 5>), dim3(grid), dim3(kBlock), 0, s, q);
  else hipLaunchKernelGGL((portfolio_adjoint_kernel<VARIANT, 0>), dim3(grid), dim3(kBlock), 0, s, q);
}

}  // namespace

int portfolio_adjoint_occupancy(int variant, int waves) {
  switch (variant) {
    case APV_GBM: return adjoint_occupancy_of<APV_GBM>(waves);
    case APV_HESTON: return adjoint_occupancy_of<APV_HESTON>(waves);
    case APV_HESTON_FEEDS: return adjoint_occupancy_of<APV_HESTON_FEEDS>(waves);
    case APV_BASKET5: return adjoint_occupancy_of<APV_BASKET5>(waves);
    case APV_HESTON_T: return adjoint_occupancy_of<APV_HESTON_T>(waves);
    case APV_HESTON_FEEDS_T: return adjoint_occupancy_of<APV_HESTON_FEEDS_T>(waves);
  }
  return 1;
}

hipError_t launch_portfolio_adjoint(int variant, const AdjPortfolioArgs& q, int grid, hipStream_t s, int waves) {
  if (grid <= 0) return hipErrorInvalidValue;
  switch (variant) {
    case APV_GBM: launch_adjoint_variant<APV_GBM>(q, grid, s, waves); break;
    case APV_HESTON: launch_adjoint_variant<APV_HESTON>(q, grid, s, waves); break;
    case APV_HESTON_FEEDS: launch_adjoint_variant<APV_HESTON_FEEDS>(q, grid, s, waves); break;
    case APV_BASKET5: launch_adjoint_variant<APV_BASKET5>(q, grid, s, waves); break;
    case APV_HESTON_T: launch_adjoint_variant<APV_HESTON_T>(q, grid, s, waves); break;
    case APV_HESTON_FEEDS_T: launch_adjoint_variant<APV_HESTON_FEEDS_T>(q, grid, s, waves); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_portfolio_adjoint_finalize(const AdjArgs* jobs, const int* n_blocks, const int* n_out, int n_jobs,
                                             KStats* out, hipStream_t s) {
  if (n_jobs <= 0) return hipSuccess;
  hipLaunchKernelGGL(portfolio_adjoint_finalize_kernel, dim3(n_jobs * kAdjMaxOut), dim3(256), 0, s, jobs, n_blocks,
                     n_out, out);
  return hipGetLastError();
}

hipError_t launch_adjoint(int model, int k, int ns, const AdjArgs& a, int n_blocks, hipStream_t s) {
#define KT_ADJ(KERN) hipLaunchKernelGGL((KERN), dim3(n_blocks), dim3(kBlock), 0, s, a)
  if (model == MK_GBM) {
    if (ns == 16) KT_ADJ((adjoint_gbm_kernel<1, 16>));
    else KT_ADJ((adjoint_gbm_kernel<1, 32>));
  } else if (model == MK_HESTON && a.tangent) {
    const bool feeds = a.n_acc > 0;
    if (ns == 16) {
      if (feeds) KT_ADJ((tangent_heston_kernel<16, true>));
      else KT_ADJ((tangent_heston_kernel<16, false>));
    } else {
      if (feeds) KT_ADJ((tangent_heston_kernel<32, true>));
      else KT_ADJ((tangent_heston_kernel<32, false>));
    }
  } else if (model == MK_HESTON) {
    const bool feeds = a.n_acc > 0;
    if (a.seg != 16 && a.seg != 32) return hipErrorInvalidValue;
    const int v = (ns == 16 ? 0 : 4) + (a.seg == 16 ? 0 : 2) + (feeds ? 1 : 0);
    switch (v) {
      case 0: KT_ADJ((adjoint_heston_kernel<16, 16, false>)); break;
      case 1: KT_ADJ((adjoint_heston_kernel<16, 16, true>)); break;
      case 2: KT_ADJ((adjoint_heston_kernel<16, 32, false>)); break;
      case 3: KT_ADJ((adjoint_heston_kernel<16, 32, true>)); break;
      case 4: KT_ADJ((adjoint_heston_kernel<32, 16, false>)); break;
      case 5: KT_ADJ((adjoint_heston_kernel<32, 16, true>)); break;
      case 6: KT_ADJ((adjoint_heston_kernel<32, 32, false>)); break;
      default: KT_ADJ((adjoint_heston_kernel<32, 32, true>)); break;
    }
  } else if (ns == 16) {
    switch (k) {
      case 2: KT_ADJ((adjoint_gbm_kernel<2, 16>)); break;
      case 3: KT_ADJ((adjoint_gbm_kernel<3, 16>)); break;
      case 4: KT_ADJ((adjoint_gbm_kernel<4, 16>)); break;
      case 5: KT_ADJ((adjoint_gbm_kernel<5, 16>)); break;
      default: return hipErrorInvalidValue;
    }
  } else {
    switch (k) {
      case 2: KT_ADJ((adjoint_gbm_kernel<2, 32>)); break;
      case 3: KT_ADJ((adjoint_gbm_kernel<3, 32>)); break;
      case 4: KT_ADJ((adjoint_gbm_kernel<4, 32>)); break;
      case 5: KT_ADJ((adjoint_gbm_kernel<5, 32>)); break;
      case 6: KT_ADJ((adjoint_gbm_kernel<6, 32>)); break;
      case 7: KT_ADJ((adjoint_gbm_kernel<7, 32>)); break;
      case 8: KT_ADJ((adjoint_gbm_kernel<8, 32>)); break;
      default: return hipErrorInvalidValue;
    }
  }
#undef KT_ADJ
  return hipGetLastError();
}

}  // namespace kt
