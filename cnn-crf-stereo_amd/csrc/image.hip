// Input preparation and decode kernels: mirror pad + normalisation, rgb->gray,
// edge weights, argmin, label remap.
#include <cmath>
#include <vector>

#include "common.hpp"

namespace ccs {

__device__ __forceinline__ int reflect(int i, int n) {
  if (n == 1) return 0;
  const int period = 2 * n;
  i %= period;
  if (i < 0) i += period;
  return i < n ? i : period - 1 - i;
}


// Pass 1: out[c][y][x] = img[c][reflect(y-p)][reflect(x-p)] (stereo/stereo.cu:879-889)
// and per-block partial sums of the padded values and their squares in float64.
// Grid (kMomentBlocks, C), 256 threads; block b takes padded rows b, b + 128, ...
__global__ __launch_bounds__(256) void k_pad_moments(const float* __restrict__ img, int H, int W, int p,
                                                     float* __restrict__ out, double* __restrict__ part) {
  const int Wo = W + 2 * p, Ho = H + 2 * p;
  const int c = blockIdx.y;
  const float* ic = img + (long)c * H * W;
  float* oc = out + (long)c * Ho * Wo;
  double s = 0.0, q = 0.0;
  for (int y = blockIdx.x; y < Ho; y += gridDim.x) {
    const float* src = ic + (long)reflect(y - p, H) * W;
    float* dst = oc + (long)y * Wo;
    for (int x = threadIdx.x; x < Wo; x += blockDim.x) {
      const float v = src[reflect(x - p, W)];
      dst[x] = v;
      s += (double)v;
      q += (double)v * (double)v;
    }
  }
  __shared__ double ss[256], sq[256];
  ss[threadIdx.x] = s;
  sq[threadIdx.x] = q;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      ss[threadIdx.x] += ss[threadIdx.x + o];
      sq[threadIdx.x] += sq[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * ((long)c * gridDim.x + blockIdx.x)] = ss[0];
    part[2 * ((long)c * gridDim.x + blockIdx.x) + 1] = sq[0];
  }
}

// Pass 2: every block sums the kMomentBlocks partials of its channel in the same
// order (so all blocks agree bit for bit), forms mean = sum / N and the
// population std = sqrt(sum x^2 / N - mean^2) (stereo/utils.cu:160-237), block 0
// publishes them, and the block normalises its share: x = (x - mean) / std.
__global__ __launch_bounds__(256) void k_normalize(float* __restrict__ x, long n, const double* __restrict__ part,
                                                   double* __restrict__ stats, int C) {
  const int c = blockIdx.y;
  __shared__ double ss[kMomentBlocks], sq[kMomentBlocks];
  if (threadIdx.x < kMomentBlocks) {
    ss[threadIdx.x] = part[2 * ((long)c * kMomentBlocks + threadIdx.x)];
    sq[threadIdx.x] = part[2 * ((long)c * kMomentBlocks + threadIdx.x) + 1];
  }
  __syncthreads();
  for (int o = kMomentBlocks / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      ss[threadIdx.x] += ss[threadIdx.x + o];
      sq[threadIdx.x] += sq[threadIdx.x + o];
    }
    __syncthreads();
  }
  const double mean = ss[0] / (double)n;
  const double var = sq[0] / (double)n - mean * mean;
  const double sd = var > 0.0 ? sqrt(var) : 0.0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    stats[c] = mean;
    stats[C + c] = sd;
  }
  const float inv = sd > 0.0 ? (float)(1.0 / sd) : 1.0f;
  const float m = (float)mean;
  float* xc = x + (long)c * n;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    xc[i] = (xc[i] - m) * inv;
}

__global__ void k_rgb2gray(const float* __restrict__ rgb, long n, float* __restrict__ gray) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  gray[i] = 0.2989f * rgb[i] + 0.5870f * rgb[n + i] + 0.1140f * rgb[2 * n + i];
}

constexpr int kMaxGaussTaps = 33;
struct GaussTaps {
  float k[kMaxGaussTaps];
  int r;
};

__global__ void k_gauss_h(const float* __restrict__ in, int H, int W, GaussTaps t, float* __restrict__ out) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y;
  if (x >= W) return;
  float a = 0.f;
  for (int i = -t.r; i <= t.r; ++i) a += t.k[i + t.r] * in[(long)y * W + reflect(x + i, W)];
  out[(long)y * W + x] = a;
}

__device__ __forceinline__ float gauss_v_at(const float* __restrict__ tmp, int H, int W, const GaussTaps& t,
                                            int x, int y) {
  float a = 0.f;
  for (int i = -t.r; i <= t.r; ++i) a += t.k[i + t.r] * tmp[(long)reflect(y + i, H) * W + x];
  return a;
}

__device__ __forceinline__ float edge_fn(float d, float alpha, float beta, float lambda) {
  return alpha > 0.f ? fmaxf(lambda * expf(-alpha * powf(d, beta)), 1e-6f) : lambda * d;
}

// Vertical blur + forward differences (stereo/stereo.cu:471-496, 525-550).
__global__ void k_edge(const float* __restrict__ tmp, int H, int W, GaussTaps t, float alpha, float beta,
                       float lambda, float* __restrict__ wx, float* __restrict__ wy) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y;
  if (x >= W) return;
  const float g = gauss_v_at(tmp, H, W, t, x, y);
  const float gx = gauss_v_at(tmp, H, W, t, reflect(x + 1, W), y);
  const float gy = gauss_v_at(tmp, H, W, t, x, reflect(y + 1, H));
  wx[(long)y * W + x] = edge_fn(fabsf(gx - g), alpha, beta, lambda);
  wy[(long)y * W + x] = edge_fn(fabsf(gy - g), alpha, beta, lambda);
}

// First strict minimum over D (ties -> lowest index), G lanes per pixel.
template <int G>
__global__ void k_argmin(const float* __restrict__ vol, long HW, int D, float dmin, float dstep,
                         float* __restrict__ out) {
  const int lane = threadIdx.x % G;
  const long p = ((long)blockIdx.x * blockDim.x + threadIdx.x) / G;
  const bool valid = p < HW;
  float best = INFINITY;
  int bi = 0x7fffffff;
  if (valid) {
    const float* v = vol + p * D;
    for (int k = lane; k < D; k += G) {
      const float c = v[k];
      if (c < best) {
        best = c;
        bi = k;
      }
    }
  }
  for (int o = G / 2; o > 0; o >>= 1) {
    const float ob = __shfl_xor(best, o, G);
    const int oi = __shfl_xor(bi, o, G);
    if (ob < best || (ob == best && oi < bi)) {
      best = ob;
      bi = oi;
    }
  }
  if (valid && lane == 0) out[p] = dmin + (float)(bi == 0x7fffffff ? 0 : bi) * dstep;
}

__global__ void k_remap(const int* __restrict__ lab, long n, float dmin, float dstep, float* __restrict__ out) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (float)lab[i] * dstep + dmin;
}

// QuadDirect sub-pixel refinement (stereo/stereo.cu:924-942, 1172-1195): the
// parabola through the three volume values around the pixel's label
// z = (int)((d - disp_min) / disp_step), with h = disp_step:
//   g = (c[z-1] - c[z+1]) / (2h),  k = max(1e-3, (c[z-1] + c[z+1] - 2 c[z]) / h^2),
//   d += g / k.
// Pixels with z <= 0 or z >= D - 1 keep d (the reference skips z = 0; at
// z = D - 1 it would read one label past the pixel).
__global__ void k_quad_fit(float* __restrict__ disp, const float* __restrict__ vol, long HW, int D, float dmin,
                           float dstep) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= HW) return;
  const float d = disp[i];
  const int z = (int)((d - dmin) / dstep);
  if (z <= 0 || z >= D - 1) return;
  const float* c = vol + i * D + z;
  const float g = (c[-1] - c[1]) / (dstep * 2.f);
  const float k = fmaxf(1e-3f, (c[-1] + c[1] - 2.f * c[0]) / (dstep * dstep));
  disp[i] = d + g / k;
}

// Left-right consistency (stereo/stereo.cu:803-817, 1198-1225): out = I1 where
// |I1 + R(x + 0.5 - I1)| <= th, else -1, I1 the left disparity and R the right
// one (negative disparities), sampled as the reference's linear, clamped
// texture: at texel coordinate u, R = (1 - a) R[i] + a R[i + 1] with
// i = floor(u - 0.5), a = u - 0.5 - i and indices clamped to the row.  (The
// texture unit rounds a to 8 fractional bits; here it is exact.)
__global__ void k_lr_check(const float* __restrict__ left, const float* __restrict__ right, int H, int W, float th,
                           float* __restrict__ out) {
  const long n = (long)H * W;
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int y = (int)(i / W), x = (int)(i - (long)y * W);
  const float I1 = left[i];
  const float u = ((float)x + 0.5f - I1) - 0.5f;
  const float fi = floorf(u);
  const float a = u - fi;
  const int i0 = (int)fminf(fmaxf(fi, 0.f), (float)(W - 1));
  const int i1 = (int)fminf(fmaxf(fi + 1.f, 0.f), (float)(W - 1));
  const float* r = right + (long)y * W;
  const float I2 = -((1.f - a) * r[i0] + a * r[i1]);
  out[i] = fabsf(I1 - I2) <= th ? I1 : -1.f;
}

}  // namespace ccs

using namespace ccs;

extern "C" int ccs_pad_normalize(ccs_handle_t h, const float* img, int C, int H, int W, int pad, float* out,
                                 double* stats) {
  CCS_CHECK(h && img && out, CCS_ERR_ARG, "ccs_pad_normalize: NULL argument");
  CCS_CHECK(C >= 1 && C <= 64 && H >= 1 && W >= 1 && pad >= 0, CCS_ERR_ARG,
            "ccs_pad_normalize: bad shape C=%d H=%d W=%d pad=%d", C, H, W, pad);
  const int Ho = H + 2 * pad, Wo = W + 2 * pad;
  const long n = (long)Ho * Wo;
  const int nb = kMomentBlocks;
  int rc = h->misc.reserve(sizeof(double) * (2 * (size_t)C * nb + 2 * C), h->stream);
  if (rc) return rc;
  double* part = (double*)h->misc.ptr;
  double* st = stats ? stats : part + 2 * (size_t)C * nb;
  k_pad_moments<<<dim3(nb, C), 256, 0, h->stream>>>(img, H, W, pad, out, part);
  k_normalize<<<dim3(256, C), 256, 0, h->stream>>>(out, n, part, st, C);
  CCS_LAUNCH_CHECK();
  return CCS_OK;
}

extern "C" int ccs_rgb_to_gray(ccs_handle_t h, const float* rgb, int H, int W, float* gray) {
  CCS_CHECK(h && rgb && gray && H > 0 && W > 0, CCS_ERR_ARG, "ccs_rgb_to_gray: bad argument");
  const long n = (long)H * W;
  k_rgb2gray<<<ceil_div(n, 256), 256, 0, h->stream>>>(rgb, n, gray);
  CCS_LAUNCH_CHECK();
  return CCS_OK;
}

extern "C" int ccs_edge_weights(ccs_handle_t h, const float* gray, int H, int W, float sigma, float alpha,
                                float beta, float lambda, float* wx, float* wy) {
  CCS_CHECK(h && gray && wx && wy && H > 0 && W > 0, CCS_ERR_ARG, "ccs_edge_weights: bad argument");
  CCS_CHECK(sigma > 0.f, CCS_ERR_ARG, "ccs_edge_weights: sigma must be positive");
  GaussTaps t;
  t.r = (int)std::ceil(3.0 * sigma);
  CCS_CHECK(2 * t.r + 1 <= kMaxGaussTaps, CCS_ERR_UNSUPPORTED, "ccs_edge_weights: sigma %g too large", sigma);
  double s = 0.0;
  std::vector<double> k(2 * t.r + 1);
  for (int i = -t.r; i <= t.r; ++i) {
    k[i + t.r] = std::exp(-(double)(i * i) / (2.0 * (double)sigma * (double)sigma));
    s += k[i + t.r];
  }
  for (int i = 0; i < 2 * t.r + 1; ++i) t.k[i] = (float)(k[i] / s);
  int rc = h->misc2.reserve(sizeof(float) * (size_t)H * W, h->stream);
  if (rc) return rc;
  float* tmp = (float*)h->misc2.ptr;
  k_gauss_h<<<dim3(ceil_div(W, 256), H), 256, 0, h->stream>>>(gray, H, W, t, tmp);
  k_edge<<<dim3(ceil_div(W, 256), H), 256, 0, h->stream>>>(tmp, H, W, t, alpha, beta, lambda, wx, wy);
  CCS_LAUNCH_CHECK();
  return CCS_OK;
}

extern "C" int ccs_argmin(ccs_handle_t h, const float* cost, int H, int W, int D, float disp_min,
                          float disp_step, float* out) {
  CCS_CHECK(h && cost && out && H > 0 && W > 0 && D > 0, CCS_ERR_ARG, "ccs_argmin: bad argument");
  const long HW = (long)H * W;
  if (D >= 64) {
    k_argmin<64><<<ceil_div(HW * 64, 256), 256, 0, h->stream>>>(cost, HW, D, disp_min, disp_step, out);
  } else {
    k_argmin<8><<<ceil_div(HW * 8, 256), 256, 0, h->stream>>>(cost, HW, D, disp_min, disp_step, out);
  }
  CCS_LAUNCH_CHECK();
  return CCS_OK;
}

extern "C" int ccs_quad_fit_direct(ccs_handle_t h, float* disparity, const float* volume, int H, int W, int D,
                                   float disp_min, float disp_step) {
  CCS_CHECK(h && disparity && volume && H > 0 && W > 0 && D > 0, CCS_ERR_ARG, "ccs_quad_fit_direct: bad argument");
  CCS_CHECK(disp_step != 0.f, CCS_ERR_ARG, "ccs_quad_fit_direct: zero disparity step");
  const long HW = (long)H * W;
  k_quad_fit<<<ceil_div(HW, 256), 256, 0, h->stream>>>(disparity, volume, HW, D, disp_min, disp_step);
  CCS_LAUNCH_CHECK();
  return CCS_OK;
}

extern "C" int ccs_left_right_check(ccs_handle_t h, const float* left, const float* right, int H, int W, float th,
                                    float* out) {
  CCS_CHECK(h && left && right && out && H > 0 && W > 0, CCS_ERR_ARG, "ccs_left_right_check: bad argument");
  const long n = (long)H * W;
  k_lr_check<<<ceil_div(n, 256), 256, 0, h->stream>>>(left, right, H, W, th, out);
  CCS_LAUNCH_CHECK();
  return CCS_OK;
}

extern "C" int ccs_remap(ccs_handle_t h, const int* labels, int H, int W, float disp_min, float disp_step,
                         float* out) {
  CCS_CHECK(h && labels && out && H > 0 && W > 0, CCS_ERR_ARG, "ccs_remap: bad argument");
  const long n = (long)H * W;
  k_remap<<<ceil_div(n, 256), 256, 0, h->stream>>>(labels, n, disp_min, disp_step, out);
  CCS_LAUNCH_CHECK();
  return CCS_OK;
}
