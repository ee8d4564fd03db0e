// kt_device.h — device-side building blocks of the gfx950 path kernels:
// Philox4x32-10 counter RNG, FP32 inverse normal, wavefront reductions.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kt_program.h"

namespace kt {

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw: "Parallel random numbers: as easy
// as 1, 2, 3", SC'11).  Each round is two 32x32->64 products (v_mad_u64_u32)
// and four xors; the key schedule is wave-uniform.
__device__ __forceinline__ uint4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                                               uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
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
// function" (GPU Computing Gems Jade, 2011), single-precision version.  The
// argument x = 2u-1 is an integer times 2^-24 and is formed exactly.
__device__ __forceinline__ float normal_from_word(uint32_t w) {
  const int i = (int)((w >> 7) | 1u) - (1 << 24);
  const float x = (float)i * 5.9604644775390625e-08f;  // 2^-24
  const float t = (1.0f - x) * (1.0f + x);
  float lw = -0.693147180559945f * __builtin_amdgcn_logf(t);  // v_log_f32 is log2
  float p;
  if (lw < 5.0f) {
    lw -= 2.5f;
    p = 2.81022636e-08f;
    p = fmaf(p, lw, 3.43273939e-07f);
    p = fmaf(p, lw, -3.5233877e-06f);
    p = fmaf(p, lw, -4.39150654e-06f);
    p = fmaf(p, lw, 0.00021858087f);
    p = fmaf(p, lw, -0.00125372503f);
    p = fmaf(p, lw, -0.00417768164f);
    p = fmaf(p, lw, 0.246640727f);
    p = fmaf(p, lw, 1.50140941f);
  } else {
    lw = __builtin_amdgcn_sqrtf(lw) - 3.0f;
    p = -0.000200214257f;
    p = fmaf(p, lw, 0.000100950558f);
    p = fmaf(p, lw, 0.00134934322f);
    p = fmaf(p, lw, -0.00367342844f);
    p = fmaf(p, lw, 0.00573950773f);
    p = fmaf(p, lw, -0.0076224613f);
    p = fmaf(p, lw, 0.00943887047f);
    p = fmaf(p, lw, 1.00167406f);
    p = fmaf(p, lw, 2.83297682f);
  }
  return 1.41421356237309505f * p * x;
}

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

}  // namespace kt
