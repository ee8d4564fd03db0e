// kt_device.h — device-side building blocks of the gfx950 path kernels:
// Philox4x32-10 counter RNG, FP32 inverse normal, wavefront reductions.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kt_program.h"

namespace kt {

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw: "Parallel random numbers: as easy
// as 1, 2, 3", SC'11).  Each round is two 32x32->64 products (v_mad_u64_u32)
// and two three-input xors (gfx950 v_bitop3_b32, truth table 0x96); the key
// schedule is wave-uniform and lives in SGPRs.
__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
  return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
}

__device__ __forceinline__ uint4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                                               uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t n0 = xor3((uint32_t)(p1 >> 32), c1, k0);
    const uint32_t n2 = xor3((uint32_t)(p0 >> 32), c3, k1);
    c1 = (uint32_t)p1;
    c3 = (uint32_t)p0;
    c0 = n0;
    c2 = n2;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return make_uint4(c0, c1, c2, c3);
}

// Standard normal from one 32-bit word.  The word selects the midpoint of one
// of 2^24 equal cells of (0,1): u = (2*floor(w/256)+1)/2^25, and z = Phi^-1(u)
// = sqrt(2) erfinv(2u-1), with erfinv from M. Giles, "Approximating the erfinv
// function" (GPU Computing Gems Jade, 2011), single-precision version; the
// sqrt(2) is folded into the polynomial coefficients.  x = 2u-1 = (2m+1)
// 2^-24 - 1 with m = w >> 8 < 2^24 is formed exactly in one FMA: m converts
// exactly, m 2^-23 is exact and the sum is a multiple of 2^-24 below 1 in
// magnitude.  1-x^2 is one fused multiply-add (one rounding, exact
// cancellation near |x| = 1).  The tail branch (|z| > ~2.9, 0.34% of draws)
// is evaluated only when some lane of the wave needs it.
#define KT_SQ2(c) ((c) * 1.41421356237309505f)
__device__ __forceinline__ float normal_from_word(uint32_t w) {
  const float x = fmaf((float)(w >> 8), 1.1920928955078125e-07f, -0.999999940395355224609375f);  // 2^-23, -(1-2^-24)
  const float t = fmaf(-x, x, 1.0f);
  const float lg = __builtin_amdgcn_logf(t);  // log2(1 - x^2) <= 0
  const float wc = fmaf(lg, -0.693147180559945f, -2.5f);  // -ln(1-x^2) - 2.5
  float p = KT_SQ2(2.81022636e-08f);
  p = fmaf(p, wc, KT_SQ2(3.43273939e-07f));
  p = fmaf(p, wc, KT_SQ2(-3.5233877e-06f));
  p = fmaf(p, wc, KT_SQ2(-4.39150654e-06f));
  p = fmaf(p, wc, KT_SQ2(0.00021858087f));
  p = fmaf(p, wc, KT_SQ2(-0.00125372503f));
  p = fmaf(p, wc, KT_SQ2(-0.00417768164f));
  p = fmaf(p, wc, KT_SQ2(0.246640727f));
  p = fmaf(p, wc, KT_SQ2(1.50140941f));
  const bool tail = lg <= -7.21347520444482f;  // -ln(1-x^2) >= 5
  if (__builtin_amdgcn_ballot_w64(tail)) {
    // The empty volatile asm keeps the compiler from evaluating this branch
    // unconditionally and selecting afterwards.
    float lt = lg;
    __asm__ volatile("" : "+v"(lt));
    const float wt = __builtin_amdgcn_sqrtf(-0.693147180559945f * lt) - 3.0f;
    float q = KT_SQ2(-0.000200214257f);
    q = fmaf(q, wt, KT_SQ2(0.000100950558f));
    q = fmaf(q, wt, KT_SQ2(0.00134934322f));
    q = fmaf(q, wt, KT_SQ2(-0.00367342844f));
    q = fmaf(q, wt, KT_SQ2(0.00573950773f));
    q = fmaf(q, wt, KT_SQ2(-0.0076224613f));
    q = fmaf(q, wt, KT_SQ2(0.00943887047f));
    q = fmaf(q, wt, KT_SQ2(1.00167406f));
    q = fmaf(q, wt, KT_SQ2(2.83297682f));
    p = tail ? q : p;
  }
  return p * x;
}
#undef KT_SQ2

__device__ __forceinline__ float exp2_fast(float x) { return __builtin_amdgcn_exp2f(x); }

// Wavefront (64-lane) sum of a double by xor-butterfly shuffles.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Chan, Golub, LeVeque pairwise update of (n, mean, M2).
__device__ __host__ __forceinline__ KStats stats_combine(const KStats& a, const KStats& b) {
  const double n = a.n + b.n;
  if (n == 0.0) return KStats{0.0, 0.0, 0.0};
  if (a.n == 0.0) return b;
  if (b.n == 0.0) return a;
  const double d = b.mean - a.mean;
  KStats r;
  r.n = n;
  r.mean = a.mean + d * (b.n / n);
  r.m2 = a.m2 + b.m2 + d * d * (a.n * b.n / n);
  return r;
}

// Statistics of one wave's values x (live lanes only), computed from shifted
// sums: the shift is lane 0's value, so a constant output has M2 == 0 exactly.
__device__ __forceinline__ KStats wave_stats(double x, bool live) {
  const double c = __shfl(x, 0, 64);
  const double d = live ? x - c : 0.0;
  const double s1 = wave_sum(d);
  const double s2 = wave_sum(d * d);
  const double n = (double)__popcll(__ballot(live));
  KStats r{n, 0.0, 0.0};
  if (n > 0.0) {
    r.mean = c + s1 / n;
    r.m2 = fmax(s2 - s1 * (s1 / n), 0.0);
  }
  return r;
}

// Deterministic combine of per-block statistics of output j (whole block).
__device__ __forceinline__ void finalize_body(const KStats* __restrict__ partials, int n_blocks, int nout, int j,
                                              KStats* __restrict__ out) {
  KStats acc{0.0, 0.0, 0.0};
  // Thread t takes blocks t, t + 256, ... eight at a time: the eight loads
  // issue together and are combined as a pairwise tree (fixed order).
  for (int b0 = threadIdx.x; b0 < n_blocks; b0 += 8 * 256) {
    KStats q[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int b = b0 + u * 256;
      q[u] = b < n_blocks ? partials[(size_t)b * nout + j] : KStats{0.0, 0.0, 0.0};
    }
#pragma unroll
    for (int w = 1; w < 8; w *= 2)
#pragma unroll
      for (int u = 0; u < 8; u += 2 * w) q[u] = stats_combine(q[u], q[u + w]);
    acc = stats_combine(acc, q[0]);
  }
  __shared__ KStats sh[256];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] = stats_combine(sh[threadIdx.x], sh[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = sh[0];
}

}  // namespace kt
