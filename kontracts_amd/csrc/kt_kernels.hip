// kt_kernels.hip — gfx950 path kernels: simulate, interpret the contract's op
// program in registers, discount, and reduce in FP64, in one pass per path.
//
// Work layout: one path per lane, 256 lanes (4 waves) per workgroup, paths
// contiguous across the grid.  A path's state (log-spot per asset, Heston
// variance, the op program's register file, the present values) never leaves
// VGPRs.  Normals come from Philox4x32-10 keyed by the seed with counter
// (block, path_lo, path_hi, 0): normal j = k*M + c of grid point k, component
// c, is word j%4 of block j/4, so the stream does not depend on the launch
// shape.  Each group of 4 grid points consumes exactly M Philox blocks.
#include "kt_device.h"
#include "kt_launch.h"

namespace kt {

// Wave-uniform selection of a per-asset kernel argument.
template <int K>
__device__ __forceinline__ float pick_asset(const float (&arr)[kMaxAssets], uint32_t i) {
  float v = arr[0];
#pragma unroll
  for (int c = 1; c < K; ++c)
    if (i == (uint32_t)c) v = arr[c];
  return v;
}

template <int NS, int NOUT, int K>
__device__ __forceinline__ void run_ops(const PathArgs& A, int begin, int count, float (&r)[NS], double (&pv)[NOUT],
                                        const float (&x)[K], double df, bool live) {
  const uint2* __restrict__ raw = reinterpret_cast<const uint2*>(A.ops) + begin;
  for (int i = 0; i < count; ++i) {
    const uint2 w = raw[i];
    const uint32_t code = w.x & 0xffu;
    const uint32_t d = (w.x >> 8) & (NS - 1);
    const uint32_t a = (w.x >> 16) & (NS - 1);
    const uint32_t b = (w.x >> 24) & (NS - 1);
    const float imm = __uint_as_float(w.y);
    const float va = r[a];
    const float vb = r[b];
    switch (code) {
      case OP_CONST: r[d] = imm; break;
      case OP_SPOT: {
        const uint32_t asset = (w.x >> 16) & 0xffu;
        const uint32_t ai = asset < (uint32_t)K ? asset : 0u;
        r[d] = pick_asset<K>(A.s0, ai) * exp2_fast(x[ai]);
        break;
      }
      case OP_NEG: r[d] = -va; break;
      case OP_ADD: r[d] = va + vb; break;
      case OP_SUB: r[d] = va - vb; break;
      case OP_MUL: r[d] = va * vb; break;
      case OP_DIV: r[d] = va / vb; break;
      case OP_MAX: r[d] = fmaxf(va, vb); break;
      case OP_MIN: r[d] = fminf(va, vb); break;
      case OP_ADDI: r[d] = va + imm; break;
      case OP_SUBI: r[d] = va - imm; break;
      case OP_RSUBI: r[d] = imm - va; break;
      case OP_MULI: r[d] = va * imm; break;
      case OP_DIVI: r[d] = va / imm; break;
      case OP_RDIVI: r[d] = imm / va; break;
      case OP_MAXI: r[d] = fmaxf(va, imm); break;
      case OP_MINI: r[d] = fminf(va, imm); break;
      case OP_GE: r[d] = va >= vb ? 1.0f : 0.0f; break;
      case OP_GT: r[d] = va > vb ? 1.0f : 0.0f; break;
      case OP_LE: r[d] = va <= vb ? 1.0f : 0.0f; break;
      case OP_LT: r[d] = va < vb ? 1.0f : 0.0f; break;
      case OP_GEI: r[d] = va >= imm ? 1.0f : 0.0f; break;
      case OP_GTI: r[d] = va > imm ? 1.0f : 0.0f; break;
      case OP_LEI: r[d] = va <= imm ? 1.0f : 0.0f; break;
      case OP_LTI: r[d] = va < imm ? 1.0f : 0.0f; break;
      case OP_AND: r[d] = va * vb; break;
      case OP_OR: r[d] = fmaxf(va, vb); break;
      case OP_NOT: r[d] = 1.0f - va; break;
      case OP_WHEN: {
        const uint32_t dn = (d + 1) & (NS - 1);
        const float n = va * (1.0f - r[d]) * vb;
        r[d] = r[d] + n;
        r[dn] = n;
        break;
      }
      case OP_UNTIL: r[d] = fmaxf(r[d], va * vb); break;
      case OP_EMIT: {
        const uint32_t o = ((w.x >> 8) & 0xffu) & (NOUT - 1);
        pv[o] += (double)(imm * (va * vb)) * df;
        break;
      }
      case OP_RMAX: r[d] = fmaxf(r[d], va); break;
      case OP_RMIN: r[d] = fminf(r[d], va); break;
      case OP_AVG: {
        const uint32_t dc = (d + 1) & (NS - 1);
        const float y = va - r[dc];
        const float t = r[d] + y;
        r[dc] = (t - r[d]) - y;
        r[d] = t;
        break;
      }
      case OP_FLAGERR:
        if (live && va != 0.0f) atomicOr(A.err, 1u << ((uint32_t)imm & 31u));
        break;
      default: break;
    }
  }
}

// Advance one path to grid point k with that point's normals zk[0..M), then
// run the point's op list.
template <int MODEL, int K, int NS, int NOUT, bool DEBUG>
__device__ __forceinline__ void step_body(const PathArgs& A, int k, const float* zk, float (&x)[K], float& v,
                                          float (&r)[NS], double (&pv)[NOUT], bool live, uint64_t local) {
  constexpr int M = MODEL == MK_GBM ? 1 : (MODEL == MK_HESTON ? 2 : K);
  if constexpr (MODEL == MK_GBM) {
    const float2 ab = reinterpret_cast<const float2*>(A.tab)[k];
    x[0] = fmaf(ab.y, zk[0], x[0] + ab.x);
  } else if constexpr (MODEL == MK_HESTON) {
    const float* t = A.tab + kHestonStride * k;
    const float z1 = zk[0], z2 = zk[1];
    const float vp = A.heston_full_truncation ? fmaxf(v, 0.0f) : v;
    const float sv = __builtin_amdgcn_sqrtf(vp);
    x[0] = x[0] + fmaf(t[1], vp, t[0]) + t[2] * sv * z1;
    const float wv = fmaf(A.rho, z1, A.rho_perp * z2);
    const float vn = fmaf(t[3], vp, v) + fmaf(t[5] * sv, wv, t[4]);
    v = A.heston_full_truncation ? vn : fmaxf(vn, 0.0f);
  } else {
    const float2* ab = reinterpret_cast<const float2*>(A.tab) + (size_t)k * K;
#pragma unroll
    for (int i = 0; i < K; ++i) {
      float wi = 0.0f;
#pragma unroll
      for (int c = 0; c <= i; ++c) wi = fmaf(A.chol[i * K + c], zk[c], wi);
      x[i] = fmaf(ab[i].y, wi, x[i] + ab[i].x);
    }
  }
  if constexpr (DEBUG) {
    if (live && A.z_out) {
#pragma unroll
      for (int c = 0; c < M; ++c) A.z_out[((size_t)local * A.n_grid + k) * M + c] = zk[c];
    }
  }
  const KStep st = A.steps[k];
  if (st.count > 0) run_ops<NS, NOUT, K>(A, st.begin, st.count, r, pv, x, A.df[k], live);
}

template <int MODEL, int K, int NS, int NOUT, bool DEBUG>
__global__ __launch_bounds__(kBlock) void path_kernel(const PathArgs A) {
  constexpr int M = MODEL == MK_GBM ? 1 : (MODEL == MK_HESTON ? 2 : K);
  const uint64_t local = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  const bool live = local < A.n_paths;
  const uint64_t path = A.path_base + local;
  const uint32_t p_lo = (uint32_t)path, p_hi = (uint32_t)(path >> 32);

  float r[NS];
#pragma unroll
  for (int i = 0; i < NS; ++i) r[i] = 0.0f;
  double pv[NOUT];
#pragma unroll
  for (int j = 0; j < NOUT; ++j) pv[j] = 0.0;
  for (int i = 0; i < A.n_init; ++i) {
    const uint2 w = reinterpret_cast<const uint2*>(A.init)[i];
    r[(w.x >> 8) & (NS - 1)] = __uint_as_float(w.y);
  }

  float x[K];
#pragma unroll
  for (int i = 0; i < K; ++i) x[i] = 0.0f;
  float v = A.v0;

  for (int g = 0; 4 * g < A.n_grid; ++g) {
    float z[4 * M];
#pragma unroll
    for (int bk = 0; bk < M; ++bk) {
      const uint4 q = philox4x32_10((uint32_t)(g * M + bk), p_lo, p_hi, 0u, A.seed_lo, A.seed_hi);
      z[4 * bk + 0] = normal_from_word(q.x);
      z[4 * bk + 1] = normal_from_word(q.y);
      z[4 * bk + 2] = normal_from_word(q.z);
      z[4 * bk + 3] = normal_from_word(q.w);
    }
    const int kend = min(4, A.n_grid - 4 * g);
    if constexpr (MODEL == MK_BASKET) {
      // 4K normals exceed what indexed register access covers: unroll the
      // four grid points so every normal index is a compile-time constant.
#pragma unroll
      for (int s = 0; s < 4; ++s)
        if (s < kend) step_body<MODEL, K, NS, NOUT, DEBUG>(A, 4 * g + s, &z[s * M], x, v, r, pv, live, local);
    } else {
      for (int s = 0; s < kend; ++s) {
        float zs[M];
#pragma unroll
        for (int c = 0; c < M; ++c) zs[c] = z[s * M + c];
        step_body<MODEL, K, NS, NOUT, DEBUG>(A, 4 * g + s, zs, x, v, r, pv, live, local);
      }
    }
  }

  if constexpr (DEBUG) {
    if (live && A.pv_out) A.pv_out[local] = pv[0];
  }

  // Block statistics per output: wave shifted sums, then Chan across waves.
  __shared__ KStats sh[NOUT][kBlock / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int j = 0; j < NOUT; ++j) {
    const KStats st = wave_stats(pv[j], live);
    if (lane == 0) sh[j][wave] = st;
  }
  __syncthreads();
  if (threadIdx.x < NOUT) {
    KStats acc = sh[threadIdx.x][0];
#pragma unroll
    for (int wv = 1; wv < kBlock / 64; ++wv) acc = stats_combine(acc, sh[threadIdx.x][wv]);
    A.partials[(size_t)blockIdx.x * NOUT + threadIdx.x] = acc;
  }
}

// Deterministic combine of per-block statistics: block j reduces output j.
__global__ __launch_bounds__(256) void finalize_kernel(const KStats* __restrict__ partials, int n_blocks, int nout,
                                                       KStats* __restrict__ out) {
  const int j = blockIdx.x;
  KStats acc{0.0, 0.0, 0.0};
  for (int b = threadIdx.x; b < n_blocks; b += 256) acc = stats_combine(acc, partials[(size_t)b * nout + j]);
  __shared__ KStats sh[256];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] = stats_combine(sh[threadIdx.x], sh[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[j] = sh[0];
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

__global__ void debug_normals_kernel(const uint32_t* __restrict__ w, uint64_t n, float* __restrict__ out) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = normal_from_word(w[i]);
}

// ---------------------------------------------------------------------------
// Variant table and launchers.
// ---------------------------------------------------------------------------
namespace {

template <int MODEL, int K, int NS, int NOUT, bool DEBUG>
hipError_t go(const PathArgs& a, int n_blocks, hipStream_t s) {
  hipLaunchKernelGGL((path_kernel<MODEL, K, NS, NOUT, DEBUG>), dim3(n_blocks), dim3(kBlock), 0, s, a);
  return hipGetLastError();
}

template <int MODEL, int K>
hipError_t go_ns(int ns, int nout, bool debug, const PathArgs& a, int n_blocks, hipStream_t s) {
  if (debug) return go<MODEL, K, 32, 1, true>(a, n_blocks, s);
  if constexpr (MODEL != MK_BASKET) {
    if (ns == 8 && nout == 1) return go<MODEL, K, 8, 1, false>(a, n_blocks, s);
    if (ns == 16 && nout == 1) return go<MODEL, K, 16, 1, false>(a, n_blocks, s);
    if (ns == 16 && nout == 8) return go<MODEL, K, 16, 8, false>(a, n_blocks, s);
  }
  if (ns == 32 && nout == 1) return go<MODEL, K, 32, 1, false>(a, n_blocks, s);
  if (ns == 32 && nout == 8) return go<MODEL, K, 32, 8, false>(a, n_blocks, s);
  return hipErrorInvalidValue;
}

}  // namespace

bool path_variant_exists(int model, int k_assets, int n_slots, int n_outputs, bool debug, int* ns_chosen,
                         int* nout_chosen) {
  if (model == MK_BASKET ? (k_assets < 2 || k_assets > kMaxAssets) : k_assets != 1) return false;
  if (n_outputs < 1 || n_outputs > 8 || n_slots > 32) return false;
  int nout = n_outputs == 1 ? 1 : 8;
  int ns;
  if (debug) {
    if (n_outputs != 1) return false;
    ns = 32;
  } else if (model == MK_BASKET) {
    ns = 32;
  } else {
    ns = n_slots <= 8 && nout == 1 ? 8 : (n_slots <= 16 ? 16 : 32);
  }
  *ns_chosen = ns;
  *nout_chosen = nout;
  return true;
}

hipError_t launch_path(int model, int k, int ns, int nout, bool debug, const PathArgs& a, int n_blocks,
                       hipStream_t s) {
  switch (model) {
    case MK_GBM: return go_ns<MK_GBM, 1>(ns, nout, debug, a, n_blocks, s);
    case MK_HESTON: return go_ns<MK_HESTON, 1>(ns, nout, debug, a, n_blocks, s);
    case MK_BASKET:
      switch (k) {
        case 2: return go_ns<MK_BASKET, 2>(ns, nout, debug, a, n_blocks, s);
        case 3: return go_ns<MK_BASKET, 3>(ns, nout, debug, a, n_blocks, s);
        case 4: return go_ns<MK_BASKET, 4>(ns, nout, debug, a, n_blocks, s);
        case 5: return go_ns<MK_BASKET, 5>(ns, nout, debug, a, n_blocks, s);
        case 6: return go_ns<MK_BASKET, 6>(ns, nout, debug, a, n_blocks, s);
        case 7: return go_ns<MK_BASKET, 7>(ns, nout, debug, a, n_blocks, s);
        case 8: return go_ns<MK_BASKET, 8>(ns, nout, debug, a, n_blocks, s);
      }
  }
  return hipErrorInvalidValue;
}

hipError_t launch_finalize(const KStats* partials, int n_blocks, int nout, int n_used, KStats* out,
                           hipStream_t s) {
  hipLaunchKernelGGL(finalize_kernel, dim3(n_used), dim3(256), 0, s, partials, n_blocks, nout, out);
  return hipGetLastError();
}

hipError_t launch_debug_philox(const uint32_t* ctr, uint32_t k0, uint32_t k1, uint64_t n, uint32_t* out,
                               hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(debug_philox_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, ctr, k0, k1, n, out);
  return hipGetLastError();
}

hipError_t launch_debug_normals(const uint32_t* words, uint64_t n, float* out, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(debug_normals_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, words, n, out);
  return hipGetLastError();
}

}  // namespace kt
