// Shared helpers of the HIP library: status plumbing, the handle, bf16 and
// wavefront (64-lane) reductions written with DPP for gfx950.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/ccs.h"

namespace ccs {

void set_error(const char* fmt, ...);

// Records the message and returns the status (usable as `return fail(...)`).
int fail(int status, const char* fmt, ...);

// Resolves the current device's address of the fixed scratch array of
// ccs_conv2d_forward_noalloc (csrc/conv.hip); ccs_create calls it.
int noalloc_scratch_setup();

#define CCS_HIP(expr)                                                                  \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess)                                                              \
      return ::ccs::fail(CCS_ERR_HIP, "%s failed: %s (%s:%d)", #expr,                  \
                         hipGetErrorString(e_), __FILE__, __LINE__);                   \
  } while (0)

#define CCS_CHECK(cond, status, ...)                \
  do {                                              \
    if (!(cond)) return ::ccs::fail(status, __VA_ARGS__); \
  } while (0)

#define CCS_LAUNCH_CHECK()                                                             \
  do {                                                                                 \
    hipError_t e_ = hipGetLastError();                                                 \
    if (e_ != hipSuccess)                                                              \
      return ::ccs::fail(CCS_ERR_HIP, "kernel launch failed: %s (%s:%d)",              \
                         hipGetErrorString(e_), __FILE__, __LINE__);                   \
  } while (0)

// Grow-only device buffer owned by a handle, allocated and released in the
// order of the handle's stream.
struct Scratch {
  void* ptr = nullptr;
  size_t bytes = 0;
  int reserve(size_t need, hipStream_t st);
  void release(hipStream_t st);
};

}  // namespace ccs

struct ccs_handle_s {
  hipStream_t stream = nullptr;
  // Named scratch regions; each stage uses its own so results of one stage can
  // stay alive while another runs.
  ccs::Scratch conv_w;     // packed conv weights (non-packed entry point)
  ccs::Scratch crf_a;      // CRF row unaries
  ccs::Scratch crf_msg;    // CRF backward messages
  ccs::Scratch crf_small;  // per-chain doubles
  ccs::Scratch crf_sums;   // per-half-step sums of ccs_slackprop
  ccs::Scratch misc;       // small temporaries
  ccs::Scratch misc2;
  ccs::Scratch conv_y;     // recomputed layer output of ccs_conv2d_backward
  ccs::Scratch crf_nan;    // per row chain: NaN seen by the first row sweep of the last run
  int nan_chains = 0;      // rows flagged in crf_nan by that sweep (0: none ran)
};

namespace ccs {

inline int feature_stride(int C) { return (C + 7) / 8 * 8; }

// Blocks of the image-moment pass (one float64 partial pair per block and
// channel) and of the float64 energy / loss reductions; ccs_reserve sizes the
// handle's scratch from these.
constexpr int kMomentBlocks = 128;
constexpr int kReduceBlocks = 256;
constexpr int kXentBlocks = 4096;

inline unsigned ceil_div(long a, long b) { return (unsigned)((a + b - 1) / b); }

typedef __bf16 bf16;
typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(4))) float f32x4;
// Native 16-byte vector (unlike HIP's uint4 struct it stays in registers when copied).
typedef __attribute__((ext_vector_type(4))) unsigned int u32x4;
typedef __attribute__((ext_vector_type(2))) unsigned int u32x2;

__device__ __forceinline__ float bf16_to_f32(bf16 v) { return (float)v; }
__device__ __forceinline__ bf16 f32_to_bf16(float v) { return (bf16)v; }

// tanh in float32 without branches: for |x| < 0.625 the odd polynomial
// x + x^3 q(x^2), q of degree 4 fitted (near-minimax, relative error) to
// (tanh(x) - x) / x^3 on that range, 0.8 ulp in float32 arithmetic; above it
// 1 - 2e / (1 + e) with e = exp(-2|x|) from v_exp_f32 and v_rcp_f32 (1.5 ulp
// with exact exp2 and reciprocal; each hardware approximation adds about one).
// Both branches are evaluated and selected, so a wave never diverges; libm's
// tanhf branches on |x| and runs both sides for mixed waves.
__device__ __forceinline__ float tanh_fast(float x) {
  const float ax = fabsf(x);
  const float s = x * x;
  float q = -0.005717044696211815f;
  q = fmaf(q, s, 0.020650586113333702f);
  q = fmaf(q, s, -0.05374349653720856f);
  q = fmaf(q, s, 0.13331490755081177f);
  q = fmaf(q, s, -0.3333328366279602f);
  const float p = fmaf(ax * s, q, ax);
  const float e = __builtin_amdgcn_exp2f(ax * -2.8853900817779268f);  // exp(-2|x|)
  const float t = fmaf(-2.f * e, __builtin_amdgcn_rcpf(1.f + e), 1.f);
  return copysignf(ax < 0.625f ? p : t, x);
}

// tanh for results that are stored in bf16 (8-bit mantissa):
// 2 / (1 + exp(-2x)) - 1.  Absolute error about 2e-7 everywhere (relative
// error grows towards 0 as 1e-7 / |x|, far under bf16's 2^-9 rounding); the
// limits are exact (exp -> inf gives -1, exp -> 0 gives 1).  Two
// transcendental and three plain instructions, no branch, no sign fix-up.
__device__ __forceinline__ float tanh_for_bf16(float x) {
  const float e = __builtin_amdgcn_exp2f(x * -2.8853900817779268f);  // exp(-2x)
  return fmaf(2.f, __builtin_amdgcn_rcpf(1.f + e), -1.f);
}

// ---------------------------------------------------------------------------
// DPP helpers (wave64).  Controls: quad_perm 0xb1 = [1,0,3,2], 0x4e = [2,3,0,1],
// row_shr:n = 0x110+n, row_bcast:15 = 0x142, row_bcast:31 = 0x143,
// wave_shl:1 = 0x130, wave_shr:1 = 0x138.

template <int CTRL, int ROW_MASK = 0xf, int BANK_MASK = 0xf, bool BOUND_CTRL = false>
__device__ __forceinline__ float dpp_f(float old, float src) {
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(src), CTRL,
                                                    ROW_MASK, BANK_MASK, BOUND_CTRL));
}

template <int CTRL>
__device__ __forceinline__ float dpp_rot(float x) {
  // Rotations give every lane a valid source, so bound_ctrl never applies; setting
  // it lets the compiler fuse the move into the consuming v_min/v_max/v_add.
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), CTRL, 0xf, 0xf, true));
}

// Minimum over the 64 lanes, left in every lane: rotations inside each row of
// 16 lanes (quad_perm, row_ror:4, row_ror:8), then v_permlane16_swap and
// v_permlane32_swap (gfx950) combine the rows -- no readlane, no partial masks.
__device__ __forceinline__ float wave_min(float x) {
  x = fminf(x, dpp_rot<0xb1>(x));
  x = fminf(x, dpp_rot<0x4e>(x));
  x = fminf(x, dpp_rot<0x124>(x));
  x = fminf(x, dpp_rot<0x128>(x));
  const auto p = __builtin_amdgcn_permlane16_swap(__float_as_uint(x), __float_as_uint(x), false, false);
  x = fminf(__uint_as_float(p[0]), __uint_as_float(p[1]));
  const auto q = __builtin_amdgcn_permlane32_swap(__float_as_uint(x), __float_as_uint(x), false, false);
  return fminf(__uint_as_float(q[0]), __uint_as_float(q[1]));
}

__device__ __forceinline__ float wave_max(float x) {
  x = fmaxf(x, dpp_rot<0xb1>(x));
  x = fmaxf(x, dpp_rot<0x4e>(x));
  x = fmaxf(x, dpp_rot<0x124>(x));
  x = fmaxf(x, dpp_rot<0x128>(x));
  const auto p = __builtin_amdgcn_permlane16_swap(__float_as_uint(x), __float_as_uint(x), false, false);
  x = fmaxf(__uint_as_float(p[0]), __uint_as_float(p[1]));
  const auto q = __builtin_amdgcn_permlane32_swap(__float_as_uint(x), __float_as_uint(x), false, false);
  return fmaxf(__uint_as_float(q[0]), __uint_as_float(q[1]));
}

__device__ __forceinline__ float wave_sum(float x) {
  x += dpp_rot<0xb1>(x);
  x += dpp_rot<0x4e>(x);
  x += dpp_rot<0x124>(x);
  x += dpp_rot<0x128>(x);
  const auto p = __builtin_amdgcn_permlane16_swap(__float_as_uint(x), __float_as_uint(x), false, false);
  x = __uint_as_float(p[0]) + __uint_as_float(p[1]);
  const auto q = __builtin_amdgcn_permlane32_swap(__float_as_uint(x), __float_as_uint(x), false, false);
  return __uint_as_float(q[0]) + __uint_as_float(q[1]);
}

// Value of the previous lane (lane 0 gets `fill`) / next lane (lane 63 gets `fill`).
__device__ __forceinline__ float lane_prev(float x, float fill) {
  return dpp_f<0x138, 0xf, 0xf, false>(fill, x);  // wave_shr:1
}
__device__ __forceinline__ float lane_next(float x, float fill) {
  return dpp_f<0x130, 0xf, 0xf, false>(fill, x);  // wave_shl:1
}

// Global -> LDS staging of n 16-byte vectors: each thread issues BATCH loads
// before it stores any of them, so a tile costs one memory latency per batch
// instead of one per vector.  load(v) is called with v clamped to [0, n) and
// returns the vector (zero where the source is outside the image); store(v, x)
// runs for v < n only.
template <int BATCH, typename Load, typename Store>
__device__ __forceinline__ void batched_copy16(int n, int tid, int nthreads, Load&& load, Store&& store) {
  for (int base = 0; base < n; base += BATCH * nthreads) {
    u32x4 r[BATCH];
#pragma unroll
    for (int i = 0; i < BATCH; ++i) r[i] = load(min(base + tid + i * nthreads, n - 1));
#pragma unroll
    for (int i = 0; i < BATCH; ++i) {
      const int v = base + tid + i * nthreads;
      if (v < n) store(v, r[i]);
    }
  }
}

__device__ __forceinline__ double wave_sum_f64(double x) {
  for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
  return x;
}

}  // namespace ccs
