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

// Grow-only device buffer owned by a handle.
struct Scratch {
  void* ptr = nullptr;
  size_t bytes = 0;
  int reserve(size_t need);
  void release();
};

}  // namespace ccs

struct ccs_handle_s {
  hipStream_t stream = nullptr;
  // Named scratch regions; each stage uses its own so results of one stage can
  // stay alive while another runs.
  ccs::Scratch conv_w;     // packed conv weights (non-packed entry point)
  ccs::Scratch crf_a;      // CRF row unaries
  ccs::Scratch crf_msg;    // CRF backward messages
  ccs::Scratch crf_small;  // per-chain doubles, reductions
  ccs::Scratch misc;       // small temporaries
  ccs::Scratch misc2;
};

namespace ccs {

inline int feature_stride(int C) { return (C + 7) / 8 * 8; }

inline unsigned ceil_div(long a, long b) { return (unsigned)((a + b - 1) / b); }

typedef __bf16 bf16;
typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(4))) float f32x4;

__device__ __forceinline__ float bf16_to_f32(bf16 v) { return (float)v; }
__device__ __forceinline__ bf16 f32_to_bf16(float v) { return (bf16)v; }

// ---------------------------------------------------------------------------
// DPP helpers (wave64).  Controls: quad_perm 0xb1 = [1,0,3,2], 0x4e = [2,3,0,1],
// row_shr:n = 0x110+n, row_bcast:15 = 0x142, row_bcast:31 = 0x143,
// wave_shl:1 = 0x130, wave_shr:1 = 0x138.

template <int CTRL, int ROW_MASK = 0xf, int BANK_MASK = 0xf, bool BOUND_CTRL = false>
__device__ __forceinline__ float dpp_f(float old, float src) {
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(src), CTRL,
                                                    ROW_MASK, BANK_MASK, BOUND_CTRL));
}

// Minimum over the 64 lanes, broadcast to every lane.
__device__ __forceinline__ float wave_min(float x) {
  x = fminf(x, dpp_f<0xb1>(x, x));
  x = fminf(x, dpp_f<0x4e>(x, x));
  x = fminf(x, dpp_f<0x114>(x, x));  // row_shr:4
  x = fminf(x, dpp_f<0x118>(x, x));  // row_shr:8
  x = fminf(x, dpp_f<0x142, 0xa>(x, x));  // row_bcast:15
  x = fminf(x, dpp_f<0x143, 0xc>(x, x));  // row_bcast:31
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), 63));
}

__device__ __forceinline__ float wave_max(float x) {
  x = fmaxf(x, dpp_f<0xb1>(x, x));
  x = fmaxf(x, dpp_f<0x4e>(x, x));
  x = fmaxf(x, dpp_f<0x114>(x, x));
  x = fmaxf(x, dpp_f<0x118>(x, x));
  x = fmaxf(x, dpp_f<0x142, 0xa>(x, x));
  x = fmaxf(x, dpp_f<0x143, 0xc>(x, x));
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), 63));
}

__device__ __forceinline__ float wave_sum(float x) {
  x += dpp_f<0xb1, 0xf, 0xf, true>(0.f, x);
  x += dpp_f<0x4e, 0xf, 0xf, true>(0.f, x);
  x += dpp_f<0x114, 0xf, 0xf, true>(0.f, x);
  x += dpp_f<0x118, 0xf, 0xf, true>(0.f, x);
  x += dpp_f<0x142, 0xa, 0xf, false>(0.f, x);
  x += dpp_f<0x143, 0xc, 0xf, false>(0.f, x);
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), 63));
}

// Value of the previous lane (lane 0 gets `fill`) / next lane (lane 63 gets `fill`).
__device__ __forceinline__ float lane_prev(float x, float fill) {
  return dpp_f<0x138, 0xf, 0xf, false>(fill, x);  // wave_shr:1
}
__device__ __forceinline__ float lane_next(float x, float fill) {
  return dpp_f<0x130, 0xf, 0xf, false>(fill, x);  // wave_shl:1
}

__device__ __forceinline__ double wave_sum_f64(double x) {
  for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
  return x;
}

}  // namespace ccs
