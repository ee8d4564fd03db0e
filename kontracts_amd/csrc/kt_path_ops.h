// kt_path_ops.h — device building blocks shared by the gfx950 kernels:
// constant-address-space table loads, the op-program interpreter and the
// per-grid-point model advance.  Included by each kernel translation unit.
#pragma once
#include "kt_device.h"
#include "kt_launch.h"

#define KT_CONST __attribute__((address_space(4)))

namespace kt {

template <class T>
__device__ __forceinline__ const KT_CONST T* cst(const T* p) {
  return (const KT_CONST T*)(p);
}

// Scalar-typed loads through the constant address space (wave-uniform index).
__device__ __forceinline__ uint32_t ld_u32(const void* p, size_t i) { return cst((const uint32_t*)p)[i]; }
__device__ __forceinline__ uint64_t ld_u64(const void* p, size_t i) { return cst((const uint64_t*)p)[i]; }
__device__ __forceinline__ float ld_f32(const float* p, size_t i) { return cst(p)[i]; }
__device__ __forceinline__ double ld_f64(const double* p, size_t i) { return cst(p)[i]; }

// Wave-uniform selection of a per-asset kernel argument.
// PA (below) is PathArgs, or PathArgs in the constant address space when a
// persistent kernel reads a job's arguments in place: every field access is
// then a scalar load.
template <int K, class ARR>
__device__ __forceinline__ float pick_asset(const ARR& arr, uint32_t i) {
  float v = arr[0];
#pragma unroll
  for (int c = 1; c < K; ++c)
    if (i == (uint32_t)c) v = arr[c];
  return v;
}

template <int K, int SC, class PA>
__device__ __forceinline__ float scen_s0(const PA& A, int sc, uint32_t asset) {
  if constexpr (SC > 1) return A.s0_scen[sc];
  return pick_asset<K>(A.s0, asset);
}

#define FOR_SC _Pragma("unroll") for (int sc = 0; sc < SC; ++sc)

// In the greeks_gbm scenario layout (base, S+h, S-h, sigma+-h, r+-h) the two
// spot-bumped scenarios differ from the base only by s0: they share its
// log-spot path log2(S/s0).  xsrc(sc) is the scenario whose log-spot sc reads.
template <int SC>
__device__ constexpr int xsrc(int sc) {
  return (SC == kScenarios && (sc == 1 || sc == 2)) ? 0 : sc;
}

// The op interpreter over a register file accessed as R(sc, slot): the
// caller's VGPR array, or a file in LDS.
template <int NS, int NOUT, int K, int SC, class PA, class RF>
__device__ __forceinline__ void run_ops_rf(const PA& A, int k, int begin, int count, const RF& R,
                                           double (&pv)[SC][NOUT], const float (&x)[SC][K], bool live) {
  for (int i = 0; i < count; ++i) {
    const uint64_t word = ld_u64(A.ops, (size_t)begin + i);
    const uint2 w = make_uint2((uint32_t)word, (uint32_t)(word >> 32));
    const uint32_t code = w.x & 0xffu;
    const uint32_t d = (w.x >> 8) & (NS - 1);
    const uint32_t a = (w.x >> 16) & (NS - 1);
    const uint32_t b = (w.x >> 24) & (NS - 1);
    const float imm = __uint_as_float(w.y);
    switch (code) {
      case OP_CONST: FOR_SC R(sc, d) = imm; break;
      case OP_SPOT: {
        const uint32_t asset = (w.x >> 16) & 0xffu;
        const uint32_t ai = asset < (uint32_t)K ? asset : 0u;
        FOR_SC R(sc, d) = scen_s0<K, SC>(A, sc, ai) * exp2_fast(x[xsrc<SC>(sc)][ai]);
        break;
      }
      case OP_NEG: FOR_SC R(sc, d) = -R(sc, a); break;
      case OP_ADD: FOR_SC R(sc, d) = R(sc, a) + R(sc, b); break;
      case OP_SUB: FOR_SC R(sc, d) = R(sc, a) - R(sc, b); break;
      case OP_MUL: FOR_SC R(sc, d) = R(sc, a) * R(sc, b); break;
      case OP_DIV: FOR_SC R(sc, d) = R(sc, a) / R(sc, b); break;
      case OP_MAX: FOR_SC R(sc, d) = fmaxf(R(sc, a), R(sc, b)); break;
      case OP_MIN: FOR_SC R(sc, d) = fminf(R(sc, a), R(sc, b)); break;
      case OP_ADDI: FOR_SC R(sc, d) = R(sc, a) + imm; break;
      case OP_SUBI: FOR_SC R(sc, d) = R(sc, a) - imm; break;
      case OP_RSUBI: FOR_SC R(sc, d) = imm - R(sc, a); break;
      case OP_MULI: FOR_SC R(sc, d) = R(sc, a) * imm; break;
      case OP_DIVI: FOR_SC R(sc, d) = R(sc, a) / imm; break;
      case OP_RDIVI: FOR_SC R(sc, d) = imm / R(sc, a); break;
      case OP_MAXI: FOR_SC R(sc, d) = fmaxf(R(sc, a), imm); break;
      case OP_MINI: FOR_SC R(sc, d) = fminf(R(sc, a), imm); break;
      case OP_GE: FOR_SC R(sc, d) = R(sc, a) >= R(sc, b) ? 1.0f : 0.0f; break;
      case OP_GT: FOR_SC R(sc, d) = R(sc, a) > R(sc, b) ? 1.0f : 0.0f; break;
      case OP_LE: FOR_SC R(sc, d) = R(sc, a) <= R(sc, b) ? 1.0f : 0.0f; break;
      case OP_LT: FOR_SC R(sc, d) = R(sc, a) < R(sc, b) ? 1.0f : 0.0f; break;
      case OP_GEI: FOR_SC R(sc, d) = R(sc, a) >= imm ? 1.0f : 0.0f; break;
      case OP_GTI: FOR_SC R(sc, d) = R(sc, a) > imm ? 1.0f : 0.0f; break;
      case OP_LEI: FOR_SC R(sc, d) = R(sc, a) <= imm ? 1.0f : 0.0f; break;
      case OP_LTI: FOR_SC R(sc, d) = R(sc, a) < imm ? 1.0f : 0.0f; break;
      case OP_AND: FOR_SC R(sc, d) = R(sc, a) * R(sc, b); break;
      case OP_OR: FOR_SC R(sc, d) = fmaxf(R(sc, a), R(sc, b)); break;
      case OP_NOT: FOR_SC R(sc, d) = 1.0f - R(sc, a); break;
      case OP_WHEN: {
        const uint32_t dn = (d + 1) & (NS - 1);
        FOR_SC {
          const float n = R(sc, a) * (1.0f - R(sc, d)) * R(sc, b);
          R(sc, d) = R(sc, d) + n;
          R(sc, dn) = n;
        }
        break;
      }
      case OP_UNTIL: FOR_SC R(sc, d) = fmaxf(R(sc, d), R(sc, a) * R(sc, b)); break;
      case OP_EMIT: {
        const uint32_t o = ((w.x >> 8) & 0xffu) & (NOUT - 1);
        FOR_SC {
          const double df = ld_f64(A.df, (size_t)A.df_index[sc] * A.n_grid + k);
          pv[sc][o] += (double)(imm * (R(sc, a) * R(sc, b))) * df;
        }
        break;
      }
      case OP_RMAX: FOR_SC R(sc, d) = fmaxf(R(sc, d), R(sc, a)); break;
      case OP_RMIN: FOR_SC R(sc, d) = fminf(R(sc, d), R(sc, a)); break;
      case OP_AVG: {
        const uint32_t dc = (d + 1) & (NS - 1);
        FOR_SC {
          const float y = R(sc, a) - R(sc, dc);
          const float t = R(sc, d) + y;
          R(sc, dc) = (t - R(sc, d)) - y;
          R(sc, d) = t;
        }
        break;
      }
      case OP_FLAGERR: {
        bool hit = false;
        FOR_SC hit = hit || R(sc, a) != 0.0f;
        if (live && hit) atomicOr(A.err, 1u << ((uint32_t)imm & 31u));
        break;
      }
      default: break;
    }
  }
}


// A register array: slot i of scenario sc is r[sc][i].
template <int SC, int NS>
struct RegArray {
  float (&r)[SC][NS];
  __device__ __forceinline__ float& operator()(int sc, uint32_t i) const { return r[sc][i]; }
};

// One scenario's register file in LDS: slot i of this lane is base[i * stride].
struct RegLds {
  float* base;
  int stride;
  __device__ __forceinline__ float& operator()(int, uint32_t i) const { return base[i * (uint32_t)stride]; }
};

template <int NS, int NOUT, int K, int SC, class PA>
__device__ __forceinline__ void run_ops(const PA& A, int k, int begin, int count, float (&r)[SC][NS],
                                        double (&pv)[SC][NOUT], const float (&x)[SC][K], bool live) {
  run_ops_rf<NS, NOUT, K, SC>(A, k, begin, count, RegArray<SC, NS>{r}, pv, x, live);
}

// One past the run of op-free groups that starts at group g, or g itself
// when g has ops or is the last group (which may be partial and always takes
// the per-point path): an op-free group's slot names the first group after
// its run, so the run's length costs one read.
template <class PA>
__device__ __forceinline__ int op_free_run_end(const PA& P, int g) {
  if (g + 1 >= P.n_groups || ld_u32(P.groups, (size_t)g * 4) != (uint32_t)G_NONE) return g;
  return min((int)ld_u32(P.groups, (size_t)g * 4 + 1), P.n_groups - 1);
}

// Advance every scenario of one path to grid point k with that point's
// normals zk[0..M).
// The 4 M normals of group g (M Philox blocks), words XORed with flip.
template <int M>
__device__ __forceinline__ void group_normals(int g, uint32_t p_lo, uint32_t p_hi, uint32_t seed_lo,
                                              uint32_t seed_hi, uint32_t flip, float (&z)[4 * M]) {
#pragma unroll
  for (int bk = 0; bk < M; ++bk) {
    const uint4 q = philox4x32_10((uint32_t)(g * M + bk), p_lo, p_hi, 0u, seed_lo, seed_hi);
    z[4 * bk + 0] = normal_from_word(q.x ^ flip);
    z[4 * bk + 1] = normal_from_word(q.y ^ flip);
    z[4 * bk + 2] = normal_from_word(q.z ^ flip);
    z[4 * bk + 3] = normal_from_word(q.w ^ flip);
  }
}

// Heston variance floor applied after each Euler step: 0 for the reference's
// floor scheme (src/simulator.rs:506-541, v = max(v', 0)), -inf for full
// truncation (v' kept, max(v, 0) used inside the step).  Under the floor
// scheme v >= 0 always, so max(v, 0) inside the step is v itself and one
// expression serves both schemes.
template <class PA>
__device__ __forceinline__ float heston_floor(const PA& A) {
  return A.heston_full_truncation ? -INFINITY : 0.0f;
}

// v' = max(v + e0 vp + d1 + d2 sqrt(vp) (rho z1 + rho' z2), floor), with
// vp = max(v, 0) and sv = sqrt(vp).  The variance recompute of the adjoint
// kernels calls this same expression, so recomputed paths are bit-identical.
__device__ __forceinline__ float heston_v_next(float v, float vp, float sv, float z1, float z2, float e0, float d1,
                                               float d2, float rho, float rho_perp, float lo) {
  const float wv = fmaf(rho, z1, rho_perp * z2);
  return fmaxf(fmaf(d2 * sv, wv, fmaf(e0, vp, v) + d1), lo);
}

// A wave-uniform value copied into a VGPR.  On gfx950 an FP32 VALU
// instruction that reads an SGPR operand issues at about half the rate of
// one whose operands are all VGPRs or literals (a loop of independent FMAs
// measured 3.8e13 against 6.7e13 lane-ops/s), so loop-invariant model
// coefficients are held in VGPRs for the hot loops.
__device__ __forceinline__ float vgpr_copy(float x) {
  float r;
  asm("v_mov_b32 %0, %1" : "=v"(r) : "s"(x));
  return r;
}

// One grid point's step coefficients (a table row, the Cholesky factor,
// rho, rho' and the variance floor) held in VGPRs.
template <int MODEL, int K, int SC>
struct StepRow {
  static constexpr int N = MODEL == MK_HESTON ? 6 : 2 * (MODEL == MK_GBM ? SC : K);
  static constexpr int NC = MODEL == MK_BASKET ? K * (K + 1) / 2 : 1;
  float c[N];
  float chol[NC];
  float rho, rho_perp, lo;
};

template <int MODEL, int K, int SC, class PA>
__device__ __forceinline__ StepRow<MODEL, K, SC> vgpr_row(const PA& A, int k) {
  using R = StepRow<MODEL, K, SC>;
  R r;
  const size_t stride = MODEL == MK_HESTON ? kHestonStride : R::N;
#pragma unroll
  for (int i = 0; i < R::N; ++i) r.c[i] = vgpr_copy(ld_f32(A.tab, (size_t)k * stride + i));
  if constexpr (MODEL == MK_BASKET) {
#pragma unroll
    for (int i = 0; i < K; ++i)
#pragma unroll
      for (int c = 0; c <= i; ++c) r.chol[i * (i + 1) / 2 + c] = vgpr_copy(A.chol[i * K + c]);
  }
  if constexpr (MODEL == MK_HESTON) {
    r.rho = vgpr_copy(A.rho);
    r.rho_perp = vgpr_copy(A.rho_perp);
    r.lo = vgpr_copy(heston_floor(A));
  }
  return r;
}

// advance() with the coefficients from a StepRow: the same expressions, so
// the same results.
template <int MODEL, int K, int SC>
__device__ __forceinline__ void advance_row(const StepRow<MODEL, K, SC>& R, const float* zk, float (&x)[SC][K],
                                            float& v) {
  if constexpr (MODEL == MK_GBM) {
    FOR_SC {
      if (xsrc<SC>(sc) == sc) x[sc][0] = fmaf(R.c[2 * sc + 1], zk[0], x[sc][0] + R.c[2 * sc]);
    }
  } else if constexpr (MODEL == MK_HESTON) {
    const float z1 = zk[0], z2 = zk[1];
    const float vp = fmaxf(v, 0.0f);
    const float sv = __builtin_amdgcn_sqrtf(vp);
    x[0][0] = fmaf(R.c[2] * sv, z1, fmaf(R.c[1], vp, x[0][0]) + R.c[0]);
    v = heston_v_next(v, vp, sv, z1, z2, R.c[3], R.c[4], R.c[5], R.rho, R.rho_perp, R.lo);
  } else {
#pragma unroll
    for (int i = 0; i < K; ++i) {
      float wi = 0.0f;
#pragma unroll
      for (int c = 0; c <= i; ++c) wi = fmaf(R.chol[i * (i + 1) / 2 + c], zk[c], wi);
      x[0][i] = fmaf(R.c[2 * i + 1], wi, x[0][i] + R.c[2 * i]);
    }
  }
}

template <int MODEL, int K, int SC, class PA>
__device__ __forceinline__ void advance(const PA& A, int k, const float* zk, float (&x)[SC][K], float& v) {
  if constexpr (MODEL == MK_GBM) {
    FOR_SC {
      if (xsrc<SC>(sc) == sc) {
        const float a = ld_f32(A.tab, ((size_t)k * SC + sc) * 2), b = ld_f32(A.tab, ((size_t)k * SC + sc) * 2 + 1);
        x[sc][0] = fmaf(b, zk[0], x[sc][0] + a);
      }
    }
  } else if constexpr (MODEL == MK_HESTON) {
    float t[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) t[i] = ld_f32(A.tab, (size_t)kHestonStride * k + i);
    const float z1 = zk[0], z2 = zk[1];
    const float vp = fmaxf(v, 0.0f);
    const float sv = __builtin_amdgcn_sqrtf(vp);
    x[0][0] = fmaf(t[2] * sv, z1, fmaf(t[1], vp, x[0][0]) + t[0]);
    v = heston_v_next(v, vp, sv, z1, z2, t[3], t[4], t[5], A.rho, A.rho_perp, heston_floor(A));
  } else {
#pragma unroll
    for (int i = 0; i < K; ++i) {
      float wi = 0.0f;
#pragma unroll
      for (int c = 0; c <= i; ++c) wi = fmaf(A.chol[i * K + c], zk[c], wi);
      const float a = ld_f32(A.tab, ((size_t)k * K + i) * 2), b = ld_f32(A.tab, ((size_t)k * K + i) * 2 + 1);
      x[0][i] = fmaf(b, wi, x[0][i] + a);
    }
  }
}

}  // namespace kt
