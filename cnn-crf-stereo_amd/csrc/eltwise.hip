// Stand-alone forms of the operators that the feature towers normally fuse
// into the convolution epilogue or the correlation epilogue: bias add, tanh,
// softmax over the channel axis (optionally negated), plus layout moves between
// the C4 (RGBA) images, planar NCHW tensors and the channel-last feature maps.
// All kernels stream memory once; 256-thread blocks, grid-stride loops.
#include <cmath>

#include "common.hpp"

namespace ccs {

constexpr int kThreads = 256;

inline unsigned grid_for(long n) {
  const long b = (n + kThreads - 1) / kThreads;
  return (unsigned)(b < 65536 ? (b > 0 ? b : 1) : 65536);
}

// y[p][c] += b[c] (channel-last, channel stride cs) -- stereo/operators/bias.cpp:58-66.
__global__ void k_bias_nhwc(float* __restrict__ y, long HW, int C, int cs, const float* __restrict__ b) {
  const long n = HW * C;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const long p = i / C;
    const int c = (int)(i - p * C);
    y[p * cs + c] += b[c];
  }
}

// y[c][p] += b[c] (planar).
__global__ void k_bias_nchw(float* __restrict__ y, long HW, int C, const float* __restrict__ b) {
  const long n = HW * C;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    y[i] += b[i / HW];
}

// y = tanh(y) -- stereo/operators/activation.cpp:33-40 (CUDNN_ACTIVATION_TANH).
__global__ void k_tanh(float* __restrict__ y, long n) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    y[i] = tanh_fast(y[i]);
}

// Softmax over the D channels of each pixel of a channel-last volume, one wave
// per pixel (stereo/operators/softmax.cpp:30-38, ACCURATE: max subtracted), with
// the negation of stereo/utils.cu:289-292 folded in when `negate`.
__global__ __launch_bounds__(256) void k_softmax_rows(float* __restrict__ x, long HW, int D, int negate) {
  const int lane = threadIdx.x & 63;
  const long p = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (p >= HW) return;
  float* v = x + p * D;
  float m = -__builtin_huge_valf();
  for (int k = lane; k < D; k += 64) m = fmaxf(m, v[k]);
  m = wave_max(m);
  float s = 0.f;
  for (int k = lane; k < D; k += 64) s += __expf(v[k] - m);
  s = wave_sum(s);
  const float inv = (negate ? -1.f : 1.f) / s;
  for (int k = lane; k < D; k += 64) v[k] = __expf(v[k] - m) * inv;
}

// Adjoint of cost = -softmax(S) over D channels, one wave per pixel:
// dS_j = p_j (sum_k p_k g_k - g_j) with p = -cost, g = dLoss/dcost.
// g and dS may alias (each lane reads g[k] before it writes dS[k]).
__global__ __launch_bounds__(256) void k_neg_softmax_bwd(const float* __restrict__ cost, const float* g, long HW,
                                                         int D, float* dS) {
  const int lane = threadIdx.x & 63;
  const long p = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (p >= HW) return;
  const float* c = cost + p * D;
  const float* gp = g + p * D;
  float s = 0.f;
  for (int k = lane; k < D; k += 64) s -= c[k] * gp[k];
  s = wave_sum(s);
  for (int k = lane; k < D; k += 64) dS[p * D + k] = -c[k] * (s - gp[k]);
}

// RGBA float4 image (row pitch in floats) -> planar C x H x W, first C channels.
__global__ void k_c4_to_planar(const float* __restrict__ c4, int H, int W, long pitch, int C,
                               float* __restrict__ out) {
  const long n = (long)H * W;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const long y = i / W, x = i - y * W;
    const float4 v = *reinterpret_cast<const float4*>(c4 + y * pitch + 4 * x);
    const float e[4] = {v.x, v.y, v.z, v.w};
    for (int c = 0; c < C; ++c) out[c * n + i] = e[c];
  }
}

// Channel-last (stride cs, fp32 or bf16) -> planar fp32 C x H x W.
template <typename T>
__global__ void k_nhwc_to_nchw(const T* __restrict__ in, long HW, int C, int cs, float* __restrict__ out) {
  const long n = HW * C;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const long c = i / HW, p = i - c * HW;
    out[i] = (float)in[p * cs + c];
  }
}

// Planar fp32 C x H x W -> channel-last fp32 with stride cs (pad channels zeroed).
__global__ void k_nchw_to_nhwc(const float* __restrict__ in, long HW, int C, int cs, float* __restrict__ out) {
  const long n = HW * cs;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const long p = i / cs;
    const int c = (int)(i - p * cs);
    out[i] = c < C ? in[(long)c * HW + p] : 0.f;
  }
}

// Head of the pairwise network (stereo/stereonet.cpp:164, 1x1 convolution
// without bias, then |.| at :402): out[k][p] = |sum_c w[k][c] x[p][c]| for
// k < K, written as K planes of H x W (the 1 x 2 x H x W wx/wy buffer).  One
// thread per pixel; x is channel-last with stride cs (a multiple of 8).
template <typename T, int KMAX>
__global__ __launch_bounds__(256) void k_pairwise_head(const T* __restrict__ x, long HW, int C, int cs,
                                                       const float* __restrict__ w, int K, float* __restrict__ out) {
  extern __shared__ float ws[];  // K x C
  for (int i = threadIdx.x; i < K * C; i += blockDim.x) ws[i] = w[i];
  __syncthreads();
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= HW) return;
  float acc[KMAX];
#pragma unroll
  for (int k = 0; k < KMAX; ++k) acc[k] = 0.f;
  const T* xp = x + p * cs;
  for (int c0 = 0; c0 < C; c0 += 8) {  // 8 channels = one 16-byte (bf16) or two (fp32) loads
    float v[8];
    if constexpr (sizeof(T) == 2) {
      const u32x4 t = *reinterpret_cast<const u32x4*>(xp + c0);
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        v[2 * e] = __uint_as_float(t[e] << 16);
        v[2 * e + 1] = __uint_as_float(t[e] & 0xffff0000u);
      }
    } else {
      const float4 a = *reinterpret_cast<const float4*>(xp + c0), b = *reinterpret_cast<const float4*>(xp + c0 + 4);
      v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w; v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
    }
#pragma unroll
    for (int k = 0; k < KMAX; ++k)
#pragma unroll
      for (int e = 0; e < 8; ++e)
        if (k < K && c0 + e < C) acc[k] = fmaf(ws[k * C + c0 + e], v[e], acc[k]);
  }
#pragma unroll
  for (int k = 0; k < KMAX; ++k)
    if (k < K) out[(long)k * HW + p] = fabsf(acc[k]);
}

// Position planes of the x/y-input model (stereo/colorstereonet.cpp:101-112):
// out[0][y][x] = (x - W/2) / W, out[1][y][x] = (y - H/2) / H, evaluated in
// double as the reference's float/double expression is, then rounded to float.
__global__ void k_xy_planes(int H, int W, float* __restrict__ out) {
  const long n = (long)H * W;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const int y = (int)(i / W), x = (int)(i - (long)y * W);
    out[i] = (float)(((double)(float)x - (double)(float)W / 2.0) / (double)(float)W);
    out[n + i] = (float)(((double)(float)y - (double)(float)H / 2.0) / (double)(float)H);
  }
}

// y *= s (iu::math::mulC, stereo/stereo.cpp:134-135).
__global__ void k_scale(float* __restrict__ y, long n, float s) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) y[i] *= s;
}

}  // namespace ccs

using namespace ccs;

extern "C" int ccs_pairwise_head(ccs_handle_t h, const void* x, int precision, int H, int W, int C,
                                 const float* w, int K, float* out) {
  CCS_CHECK(h && x && w && out, CCS_ERR_ARG, "ccs_pairwise_head: NULL argument");
  CCS_CHECK(H >= 1 && W >= 1 && C >= 1 && C <= 1024 && K >= 1 && K <= 4, CCS_ERR_ARG,
            "ccs_pairwise_head: bad shape (H=%d W=%d C=%d K=%d; K <= 4)", H, W, C, K);
  const long HW = (long)H * W;
  const int cs = feature_stride(C);
  const size_t lds = sizeof(float) * (size_t)K * C;
  const unsigned grid = ceil_div(HW, 256);
  if (precision == CCS_BF16)
    k_pairwise_head<bf16, 4><<<grid, 256, lds, h->stream>>>((const bf16*)x, HW, C, cs, w, K, out);
  else if (precision == CCS_FP32)
    k_pairwise_head<float, 4><<<grid, 256, lds, h->stream>>>((const float*)x, HW, C, cs, w, K, out);
  else
    return fail(CCS_ERR_ARG, "ccs_pairwise_head: unknown precision %d", precision);
  CCS_LAUNCH_CHECK();
  return CCS_OK;
}

extern "C" int ccs_xy_planes(ccs_handle_t h, int H, int W, float* out) {
  CCS_CHECK(h && out && H >= 1 && W >= 1, CCS_ERR_ARG, "ccs_xy_planes: bad argument");
  k_xy_planes<<<grid_for((long)H * W), kThreads, 0, h->stream>>>(H, W, out);
  CCS_LAUNCH_CHECK();
  return CCS_OK;
}

extern "C" int ccs_scale_forward(ccs_handle_t h, float* y, long count, float s) {
  CCS_CHECK(h && y && count >= 0, CCS_ERR_ARG, "ccs_scale_forward: bad argument");
  if (count == 0) return CCS_OK;
  k_scale<<<grid_for(count), kThreads, 0, h->stream>>>(y, count, s);
  CCS_LAUNCH_CHECK();
  return CCS_OK;
}

extern "C" int ccs_bias_forward(ccs_handle_t h, float* y, int layout, int C, int H, int W, int cs,
                                const float* bias) {
  CCS_CHECK(h && y && bias, CCS_ERR_ARG, "ccs_bias_forward: NULL argument");
  CCS_CHECK(C >= 1 && H >= 1 && W >= 1, CCS_ERR_ARG, "ccs_bias_forward: bad shape");
  const long HW = (long)H * W;
  if (layout == CCS_LAYOUT_NHWC) {
    CCS_CHECK(cs >= C, CCS_ERR_ARG, "ccs_bias_forward: channel stride %d < C=%d", cs, C);
    k_bias_nhwc<<<grid_for(HW * C), kThreads, 0, h->stream>>>(y, HW, C, cs, bias);
  } else {
    CCS_CHECK(layout == CCS_LAYOUT_NCHW, CCS_ERR_ARG, "ccs_bias_forward: unknown layout %d", layout);
    k_bias_nchw<<<grid_for(HW * C), kThreads, 0, h->stream>>>(y, HW, C, bias);
  }
  CCS_LAUNCH_CHECK();
  return CCS_OK;
}

extern "C" int ccs_tanh_forward(ccs_handle_t h, float* y, long count) {
  CCS_CHECK(h && y && count >= 0, CCS_ERR_ARG, "ccs_tanh_forward: bad argument");
  if (count == 0) return CCS_OK;
  k_tanh<<<grid_for(count), kThreads, 0, h->stream>>>(y, count);
  CCS_LAUNCH_CHECK();
  return CCS_OK;
}

extern "C" int ccs_softmax_forward(ccs_handle_t h, float* x, long HW, int D, int negate) {
  CCS_CHECK(h && x && HW >= 0 && D >= 1, CCS_ERR_ARG, "ccs_softmax_forward: bad argument");
  if (HW == 0) return CCS_OK;
  k_softmax_rows<<<(unsigned)((HW + 3) / 4), 256, 0, h->stream>>>(x, HW, D, negate);
  CCS_LAUNCH_CHECK();
  return CCS_OK;
}

extern "C" int ccs_neg_softmax_backward(ccs_handle_t h, const float* cost, const float* dcost, long HW, int D,
                                        float* dS) {
  CCS_CHECK(h && cost && dcost && dS && HW >= 0 && D >= 1, CCS_ERR_ARG, "ccs_neg_softmax_backward: bad argument");
  if (HW == 0) return CCS_OK;
  k_neg_softmax_bwd<<<(unsigned)((HW + 3) / 4), 256, 0, h->stream>>>(cost, dcost, HW, D, dS);
  CCS_LAUNCH_CHECK();
  return CCS_OK;
}

extern "C" int ccs_c4_to_planar(ccs_handle_t h, const float* c4, int H, int W, long pitch, int C,
                                float* out) {
  CCS_CHECK(h && c4 && out, CCS_ERR_ARG, "ccs_c4_to_planar: NULL argument");
  CCS_CHECK(H >= 1 && W >= 1 && C >= 1 && C <= 4 && pitch >= 4L * W && pitch % 4 == 0, CCS_ERR_ARG,
            "ccs_c4_to_planar: bad shape (H=%d W=%d C=%d pitch=%ld)", H, W, C, pitch);
  CCS_CHECK(((uintptr_t)c4 & 15) == 0, CCS_ERR_ARG, "ccs_c4_to_planar: image must be 16-byte aligned");
  k_c4_to_planar<<<grid_for((long)H * W), kThreads, 0, h->stream>>>(c4, H, W, pitch, C, out);
  CCS_LAUNCH_CHECK();
  return CCS_OK;
}

extern "C" int ccs_nhwc_to_nchw(ccs_handle_t h, const void* in, int precision, int C, int cs, int H, int W,
                                float* out) {
  CCS_CHECK(h && in && out, CCS_ERR_ARG, "ccs_nhwc_to_nchw: NULL argument");
  CCS_CHECK(C >= 1 && cs >= C && H >= 1 && W >= 1, CCS_ERR_ARG, "ccs_nhwc_to_nchw: bad shape");
  const long HW = (long)H * W;
  if (precision == CCS_BF16)
    k_nhwc_to_nchw<bf16><<<grid_for(HW * C), kThreads, 0, h->stream>>>((const bf16*)in, HW, C, cs, out);
  else if (precision == CCS_FP32)
    k_nhwc_to_nchw<float><<<grid_for(HW * C), kThreads, 0, h->stream>>>((const float*)in, HW, C, cs, out);
  else
    return fail(CCS_ERR_ARG, "ccs_nhwc_to_nchw: unknown precision %d", precision);
  CCS_LAUNCH_CHECK();
  return CCS_OK;
}

extern "C" int ccs_nchw_to_nhwc(ccs_handle_t h, const float* in, int C, int cs, int H, int W, float* out) {
  CCS_CHECK(h && in && out, CCS_ERR_ARG, "ccs_nchw_to_nhwc: NULL argument");
  CCS_CHECK(C >= 1 && cs >= C && H >= 1 && W >= 1, CCS_ERR_ARG, "ccs_nchw_to_nhwc: bad shape");
  const long HW = (long)H * W;
  k_nchw_to_nhwc<<<grid_for(HW * cs), kThreads, 0, h->stream>>>(in, HW, C, cs, out);
  CCS_LAUNCH_CHECK();
  return CCS_OK;
}
