// Correlation cost volume (stereo/operators/stereocorrelation.cu:111-142) with
// the -softmax over disparities (stereo/operators/softmax.cpp:30-38,
// stereo/utils.cu:289-292) fused into its epilogue, and the adjoints.
//
//   S[h][w][k] = sum_c f0[h][w][c] f1[h][w + o_k][c],  o_k = -(disp_min + k disp_step),
//   0 where w + o_k leaves [0, W).
//
// Fast path k_corr_volume: one workgroup per (image row, 16*NW pixels).  Both
// feature rows are staged in LDS (f1 as the window of columns the band needs,
// zero outside the image), each wave computes the 16 x (16*NTM) block
// S[w][w'] of its 16 pixels against the window with MFMA (a banded GEMM over
// the channels), parks it in LDS over the dead feature tiles, and the epilogue
// gathers the band w' = w + o_k, applies -softmax over k and writes the
// H x W x D volume with 16-byte stores, D contiguous floats per pixel.
#include <algorithm>
#include <cmath>
#include <type_traits>
#include <vector>

#include "common.hpp"

namespace ccs {

template <typename T>
struct CorrTraits;
template <>
struct CorrTraits<bf16> {
  static constexpr int kChunk = 32;  // channels per MFMA
};
template <>
struct CorrTraits<float> {
  static constexpr int kChunk = 4;
};

struct CorrArgs {
  const void* f0;
  const void* f1;
  float* out;
  int H, W, Cs, D;
  int o0, ostep;  // o_k = o0 + k * ostep
  int o_lo;       // min_k o_k
  int epilogue;
  int pitch;      // LDS pixel pitch (elements)
  int win;        // f1 window columns
  int sp;         // S row pitch (floats)
  int stats_off;  // byte offset of the softmax statistics in LDS
};

template <typename T, int NTM>
__global__ __launch_bounds__(512) void k_corr_volume(CorrArgs a) {
  extern __shared__ __align__(16) unsigned char lds[];
  const int NW = blockDim.x >> 6;
  const int TW = 16 * NW;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int lrow = lane & 15, lgrp = lane >> 4;
  const int h = blockIdx.y;
  const int w0 = blockIdx.x * TW;
  const int Cs = a.Cs, P = a.pitch;
  T* F0 = reinterpret_cast<T*>(lds);
  T* F1 = F0 + TW * P;
  // S takes the feature tiles' place once every wave has finished its MFMAs.
  float* S = reinterpret_cast<float*>(lds);
  float* S_w = S + wave * 16 * a.sp;
  float* stats = reinterpret_cast<float*>(lds + a.stats_off);  // [NW][16]

  const T* f0 = reinterpret_cast<const T*>(a.f0) + (long)h * a.W * Cs;
  const T* f1 = reinterpret_cast<const T*>(a.f1) + (long)h * a.W * Cs;
  const int wbase = w0 + a.o_lo;  // first f1 column of the window
  // Both feature rows go to LDS in batches of 16-byte loads (one memory latency
  // per batch); columns outside the image read as zero.
  const int nthr = blockDim.x;
  const int vpp = Cs * (int)sizeof(T) / 16;  // 16-byte vectors per pixel
  auto fetch = [&](const T* row, int wi) -> const u32x4* {
    return reinterpret_cast<const u32x4*>(row + (long)min(max(wi, 0), a.W - 1) * Cs);
  };
  auto put = [&](T* dst, int px, int g, const u32x4& val) {
    if constexpr (sizeof(T) == 2) {
      *reinterpret_cast<u32x4*>(dst + px * P + 8 * g) = val;
    } else {  // odd float pitch: scalar stores
#pragma unroll
      for (int e = 0; e < 4; ++e) dst[px * P + 4 * g + e] = __uint_as_float(val[e]);
    }
  };
  batched_copy16<8>(
      TW * vpp, tid, nthr,
      [&](int v) {
        const int px = v / vpp, g = v % vpp, wi = w0 + px;
        const u32x4 val = fetch(f0, wi)[g];
        return wi < a.W ? val : u32x4{0u, 0u, 0u, 0u};
      },
      [&](int v, const u32x4& val) { put(F0, v / vpp, v % vpp, val); });
  batched_copy16<8>(
      a.win * vpp, tid, nthr,
      [&](int v) {
        const int px = v / vpp, g = v % vpp, wi = wbase + px;
        const u32x4 val = fetch(f1, wi)[g];
        return (wi >= 0 && wi < a.W) ? val : u32x4{0u, 0u, 0u, 0u};
      },
      [&](int v, const u32x4& val) { put(F1, v / vpp, v % vpp, val); });
  __syncthreads();

  f32x4 acc[NTM];
#pragma unroll
  for (int t = 0; t < NTM; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int arow = 16 * wave + lrow;  // pixel of this lane's A row
  if constexpr (sizeof(T) == 2) {
    const int cg = Cs / 8;
    for (int q = 0; q < (cg + 3) / 4; ++q) {
      const int g = 4 * q + lgrp;
      const bool ok = g < cg;
      const bf16x8 av = ok ? *reinterpret_cast<const bf16x8*>(F0 + arow * P + 8 * g) : bf16x8{};
#pragma unroll
      for (int t = 0; t < NTM; ++t) {
        const int col = 16 * wave + 16 * t + lrow;
        const bf16x8 bv = ok ? *reinterpret_cast<const bf16x8*>(F1 + col * P + 8 * g) : bf16x8{};
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc[t], 0, 0, 0);
      }
    }
  } else {
    for (int q = 0; q < Cs / 4; ++q) {
      const int c = 4 * q + lgrp;
      const float av = F0[arow * P + c];
#pragma unroll
      for (int t = 0; t < NTM; ++t) {
        const int col = 16 * wave + 16 * t + lrow;
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, F1[col * P + c], acc[t], 0, 0, 0);
      }
    }
  }
  __syncthreads();  // the feature tiles are dead: S overwrites them
  // S_w[r][j] = f0[pixel 16*wave + r] . f1[window column 16*wave + j]
#pragma unroll
  for (int t = 0; t < NTM; ++t)
#pragma unroll
    for (int e = 0; e < 4; ++e) S_w[(4 * lgrp + e) * a.sp + 16 * t + lrow] = acc[t][e];
  __syncthreads();

  // Band entry of pixel r, disparity k: S_w[r * sp + r + o_k - o_lo] = band(r)[k * ostep].
  const int D = a.D;
  const bool neg_softmax = a.epilogue == CCS_EPILOGUE_NEG_SOFTMAX;
  auto band = [&](int r) { return S_w + r * a.sp + r + a.o0 - a.o_lo; };
  if (neg_softmax) {
    // 4 lanes per pixel: max, then exp(s - max) written back in place, then 1 / sum.
    const int r = lane >> 2, s4 = lane & 3;
    float* br = band(r);
    float mx = -INFINITY;
    for (int k = s4; k < D; k += 4) mx = fmaxf(mx, br[k * a.ostep]);
    mx = fmaxf(mx, __shfl_xor(mx, 1, 64));
    mx = fmaxf(mx, __shfl_xor(mx, 2, 64));
    float sum = 0.f;
    for (int k = s4; k < D; k += 4) {
      const float e = __expf(br[k * a.ostep] - mx);
      br[k * a.ostep] = e;
      sum += e;
    }
    sum += __shfl_xor(sum, 1, 64);
    sum += __shfl_xor(sum, 2, 64);
    if (s4 == 0) stats[wave * 16 + r] = -1.f / sum;
  }
  __syncthreads();
  const long pix0 = (long)h * a.W + w0 + 16 * wave;
  const int npx = min(16, a.W - (w0 + 16 * wave));
  if (npx <= 0) return;
  float* outp = a.out + pix0 * D;
  const int lp = D >> 2;  // lanes per pixel when every lane stores 4 disparities
  if ((D & 3) == 0 && lp <= 64 && (64 % lp) == 0) {
    const int ppi = 64 / lp, rr = lane / lp, k = 4 * (lane % lp);
    for (int r0 = 0; r0 < npx; r0 += ppi) {
      const int r = r0 + rr;
      if (r < npx) {
        const float* br = band(r);
        const float sc = neg_softmax ? stats[wave * 16 + r] : 1.f;
        float4 v;
        v.x = br[k * a.ostep] * sc;
        v.y = br[(k + 1) * a.ostep] * sc;
        v.z = br[(k + 2) * a.ostep] * sc;
        v.w = br[(k + 3) * a.ostep] * sc;
        *reinterpret_cast<float4*>(outp + (long)r * D + k) = v;
      }
    }
  } else {
    for (int f = lane; f < npx * D; f += 64) {
      const int r = f / D, k = f % D;
      outp[f] = band(r)[k * a.ostep] * (neg_softmax ? stats[wave * 16 + r] : 1.f);
    }
  }
}

// Reference-compatible correlation on N x C x H x W float32 inputs (any offsets,
// rect_corr); output NHWC or NCHW.  One thread per output element.
__global__ void k_corr_generic(const float* __restrict__ im0, const float* __restrict__ im1,
                               const int* __restrict__ disp, float* __restrict__ out, int N, int C, int H, int W,
                               int D, int nhwc, int rc) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long total = (long)N * H * W * D;
  if (idx >= total) return;
  int n, h, w, k;
  if (nhwc) {
    k = (int)(idx % D);
    w = (int)((idx / D) % W);
    h = (int)((idx / ((long)D * W)) % H);
    n = (int)(idx / ((long)D * W * H));
  } else {
    w = (int)(idx % W);
    h = (int)((idx / W) % H);
    k = (int)((idx / ((long)W * H)) % D);
    n = (int)(idx / ((long)W * H * D));
  }
  const long HW = (long)H * W;
  const int w2 = w + disp[k];
  float best = -100000.f;
  for (int yy = -rc; yy <= rc; ++yy) {
    const int h2 = h + yy;
    if (h2 < 0 || h2 >= H) continue;
    float r = 0.f;
    if (w2 >= 0 && w2 < W)
      for (int c = 0; c < C; ++c)
        r += im0[((long)n * C + c) * HW + (long)h * W + w] * im1[((long)n * C + c) * HW + (long)h2 * W + w2];
    if (r > best) best = r;
  }
  out[idx] = best;
}

// Reference-compatible adjoint (stereo/operators/stereocorrelation.cu:243-278):
// one thread per (n, c, h, w), gather over disparities, no atomics.
__global__ void k_corr_bwd_generic(float* __restrict__ d0, float* __restrict__ d1, const float* __restrict__ G,
                                   const float* __restrict__ im0, const float* __restrict__ im1,
                                   const int* __restrict__ disp, int N, int C, int H, int W, int D, int nhwc) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long HW = (long)H * W;
  if (idx >= (long)N * C * HW) return;
  const int w = (int)(idx % W);
  const int h = (int)((idx / W) % H);
  const int c = (int)((idx / HW) % C);
  const int n = (int)(idx / (HW * C));
  auto g_at = [&](int hh, int ww, int k) {
    return nhwc ? G[(((long)n * H + hh) * W + ww) * D + k] : G[(((long)n * D + k) * H + hh) * W + ww];
  };
  float g0 = 0.f, g1 = 0.f;
  for (int k = 0; k < D; ++k) {
    const int wp = w + disp[k];
    if (wp >= 0 && wp < W) g0 += im1[((long)n * C + c) * HW + (long)h * W + wp] * g_at(h, w, k);
    const int wm = w - disp[k];
    if (wm >= 0 && wm < W) g1 += im0[((long)n * C + c) * HW + (long)h * W + wm] * g_at(h, wm, k);
  }
  d0[idx] = g0;
  d1[idx] = g1;
}

// Adjoint of the fast path (channel-last maps, float32 or bfloat16, gradients
// in the maps' type, float32 accumulation): one thread per (h, w, c).
template <typename T>
__global__ void k_corr_volume_bwd(const T* __restrict__ f0, const T* __restrict__ f1, int H, int W, int C,
                                  int Cs, int o0, int ostep, int D, const float* __restrict__ G,
                                  T* __restrict__ df0, T* __restrict__ df1) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)H * W * Cs) return;
  const int c = (int)(idx % Cs);
  const long p = idx / Cs;
  const int w = (int)(p % W), h = (int)(p / W);
  float g0 = 0.f, g1 = 0.f;
  if (c < C) {
    const long row = (long)h * W;
    for (int k = 0; k < D; ++k) {
      const int o = o0 + k * ostep;
      const int wp = w + o;
      if (wp >= 0 && wp < W) g0 += (float)f1[(row + wp) * Cs + c] * G[(row + w) * D + k];
      const int wm = w - o;
      if (wm >= 0 && wm < W) g1 += (float)f0[(row + wm) * Cs + c] * G[(row + wm) * D + k];
    }
  }
  df0[idx] = (T)g0;
  df1[idx] = (T)g1;
}

// Adjoint of the fast path on MFMA.  Both feature gradients have the form
//   out[p][c] = sum_k Gt[p][k] F[p + e_k][c],  e_k = e0 + k estep,
// df0: Gt = G, F = f1, e_k = o_k;  df1: Gt[p][k] = G[p - o_k][k], F = f0,
// e_k = -o_k (each G entry is read once either way; terms whose source pixel
// leaves the image are zero).  For a block of 64 output pixels of one row this
// is a banded GEMM over the window of F columns: per chunk of KC disparities
// the chunk's Gt (64 x KC) and F window (64 + span columns, all channels, as
// float32) go to LDS, and each wave computes its 16 pixels x 16*CT channels
// with v_mfma_f32_16x16x4_f32: A[m][r] = Gt[m][k] where window column r lies
// at e_k from pixel m (0 off the band), B[r][n] = F[column r][16t+n].  The
// band lookup k(d) of a column offset d is a small LDS table per chunk.  No
// atomics: every output is written once by the block that owns it.
constexpr int kBwdTW = 64;  // output pixels per block (four waves of 16)
constexpr int kBwdKC = 64;  // disparities per chunk at |estep| = 1

template <typename T, int CT>
__global__ __launch_bounds__(256, 2) void k_corr_bwd_mfma(const T* __restrict__ fsrc, const float* __restrict__ G,
                                                          int H, int W, int Cs, int D, int e0, int estep,
                                                          int gather, int o0, int ostep, T* __restrict__ out) {
  constexpr int CP = 16 * CT, TW = kBwdTW;
  extern __shared__ __align__(16) unsigned char lds[];
  const int astep = estep < 0 ? -estep : estep;
  const int KC = max(1, kBwdKC / astep);
  const int span_max = (KC - 1) * astep;
  float* Gt = reinterpret_cast<float*>(lds);          // [TW][KC]
  float* Fw = Gt + TW * KC;                           // [TW + span_max][CP]
  int* ktab = reinterpret_cast<int*>(Fw + (TW + span_max) * CP);  // [TW + span_max + 16]
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int lm = lane & 15, lr = lane >> 4;
  const int h = blockIdx.y, p0 = blockIdx.x * TW;
  const long rowpix = (long)h * W;
  const T* frow = fsrc + rowpix * Cs;
  const float* grow = G + rowpix * D;
  constexpr int VE = 16 / (int)sizeof(T);
  const int vpp = Cs / VE;
  // zero the pad channels of the window once (they are never written again)
  for (int i = tid; i < (TW + span_max) * CP; i += 256) Fw[i] = 0.f;
  f32x4 acc[CT];
#pragma unroll
  for (int t = 0; t < CT; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int kc = 0; kc < D; kc += KC) {
    const int kn = min(KC, D - kc);
    const int span = (kn - 1) * astep;
    const int emin = estep > 0 ? e0 + kc * estep : e0 + (kc + kn - 1) * estep;
    const int win = TW + span;
    __syncthreads();  // the previous chunk's MFMAs are done with Gt, Fw, ktab
    // Gt[p][kk] of this chunk
    for (int i = tid; i < TW * KC; i += 256) {
      const int p = i / KC, kk = i % KC;
      float v = 0.f;
      if (kk < kn && p0 + p < W) {
        const int k = kc + kk;
        const int src = gather ? p0 + p - (o0 + k * ostep) : p0 + p;
        if (src >= 0 && src < W) v = grow[(long)src * D + k];
      }
      Gt[i] = v;
    }
    // F window: columns p0 + emin .. p0 + emin + win - 1, zero outside the image
    for (int v = tid; v < win * vpp; v += 256) {
      const int px = v / vpp, g = v % vpp, col = p0 + emin + px;
      u32x4 r = {0u, 0u, 0u, 0u};
      if (col >= 0 && col < W) r = *reinterpret_cast<const u32x4*>(frow + (long)col * Cs + g * VE);
      float* d = Fw + px * CP + g * VE;
      if constexpr (VE == 4) {
        *reinterpret_cast<u32x4*>(d) = r;
      } else {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          d[2 * e] = __uint_as_float(r[e] << 16);
          d[2 * e + 1] = __uint_as_float(r[e] & 0xffff0000u);
        }
      }
    }
    // ktab[d + 15]: the chunk-local disparity at column offset d from a pixel
    // (window column index minus the pixel's index), or -1 off the band.
    const int dbase = emin - e0 - kc * estep;  // offset of window column 0 relative to e_kc
    for (int i = tid; i < win + 16; i += 256) {
      const int dd = dbase + (i - 15);
      int kk = -1;
      if (dd % estep == 0) {
        const int q = dd / estep;
        if (q >= 0 && q < kn) kk = q;
      }
      ktab[i] = kk;
    }
    __syncthreads();
    // wave tile: pixels 16 wave + m, window columns 16 wave + j, j < 16 + span
    const int steps = (16 + span + 3) / 4;
    for (int st = 0; st < steps; ++st) {
      const int j = 16 * wave + 4 * st + lr;  // window column of this lane's A entry
      const int kk = ktab[4 * st + lr - lm + 15];
      const float av = kk >= 0 ? Gt[(16 * wave + lm) * KC + kk] : 0.f;
      const bool jin = j < win;
#pragma unroll
      for (int t = 0; t < CT; ++t) {
        const float bv = jin ? Fw[j * CP + 16 * t + lm] : 0.f;
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv, acc[t], 0, 0, 0);
      }
    }
  }
#pragma unroll
  for (int t = 0; t < CT; ++t)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int p = p0 + 16 * wave + 4 * lr + e, c = 16 * t + lm;
      if (p < W && c < Cs) out[(rowpix + p) * Cs + c] = (T)acc[t][e];
    }
}

// Cross entropy of softmax_D(S) at the target label; dS = softmax - onehot.
__global__ void k_softmax_xent(const float* __restrict__ S, long HW, int D, const int* __restrict__ target,
                               float* __restrict__ dS, double* __restrict__ part) {
  double loss = 0.0;
  const int sub = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (long p = (long)blockIdx.x * (blockDim.x >> 6) + wv; p < HW; p += (long)gridDim.x * (blockDim.x >> 6)) {
    const float* s = S + p * D;
    float mx = -INFINITY;
    for (int k = sub; k < D; k += 64) mx = fmaxf(mx, s[k]);
    for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o, 64));
    float sum = 0.f;
    for (int k = sub; k < D; k += 64) sum += expf(s[k] - mx);
    for (int o = 32; o > 0; o >>= 1) sum += __shfl_xor(sum, o, 64);
    const float lse = mx + logf(sum);
    const int t = target[p] < D ? target[p] : -1;  // labels outside [0, D) are ignored
    for (int k = sub; k < D; k += 64) dS[p * D + k] = t >= 0 ? expf(s[k] - lse) - (k == t ? 1.f : 0.f) : 0.f;
    if (sub == 0 && t >= 0) loss += (double)(lse - s[t]);
  }
  __shared__ double sh[256];
  sh[threadIdx.x] = loss;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// Deterministic sum of n doubles (one block of 256 threads, a fixed tree).
__global__ void k_sum_f64_corr(const double* __restrict__ in, int n, double* __restrict__ out) {
  __shared__ double sh[256];
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) s += in[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = sh[0];
}

// LDS bytes of one workgroup: the feature tiles, later overlaid by S, then the statistics.
static size_t corr_lds_bytes(CorrArgs& a, int NW, size_t esz) {
  const size_t feat = (size_t)(16 * NW + a.win) * a.pitch * esz;
  const size_t sbytes = (size_t)NW * 16 * a.sp * 4;
  a.stats_off = (int)((feat > sbytes ? feat : sbytes) + 15) / 16 * 16;
  return (size_t)a.stats_off + (size_t)NW * 16 * 4;
}

template <typename T, int NTM>
static int launch_volume(CorrArgs a, int NW, hipStream_t st) {
  const size_t bytes = corr_lds_bytes(a, NW, sizeof(T));
  static bool configured = false;
  if (!configured) {
    CCS_HIP(hipFuncSetAttribute((const void*)k_corr_volume<T, NTM>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                160 * 1024));
    configured = true;
  }
  k_corr_volume<T, NTM><<<dim3(ceil_div(a.W, 16 * NW), a.H), 64 * NW, bytes, st>>>(a);
  CCS_LAUNCH_CHECK();
  return CCS_OK;
}

}  // namespace ccs

using namespace ccs;

extern "C" int ccs_correlation_volume(ccs_handle_t h, const void* f0, const void* f1, int precision, int H, int W,
                                      int C, int disp_min, int disp_step, int D, int epilogue, float* out) {
  CCS_CHECK(h && f0 && f1 && out, CCS_ERR_ARG, "ccs_correlation_volume: NULL argument");
  CCS_CHECK(H >= 1 && W >= 1 && C >= 1 && D >= 1, CCS_ERR_ARG, "ccs_correlation_volume: empty shape");
  CCS_CHECK(precision == CCS_FP32 || precision == CCS_BF16, CCS_ERR_ARG, "ccs_correlation_volume: bad precision");
  CCS_CHECK(epilogue == CCS_EPILOGUE_RAW || epilogue == CCS_EPILOGUE_NEG_SOFTMAX, CCS_ERR_ARG,
            "ccs_correlation_volume: bad epilogue");
  CCS_CHECK(disp_step != 0 || D == 1, CCS_ERR_ARG, "ccs_correlation_volume: zero disparity step");
  CCS_CHECK(C <= 256, CCS_ERR_UNSUPPORTED, "ccs_correlation_volume: more than 256 channels");
  CorrArgs a{};
  a.f0 = f0; a.f1 = f1; a.out = out; a.H = H; a.W = W; a.Cs = feature_stride(C); a.D = D;
  a.o0 = -disp_min; a.ostep = -disp_step; a.epilogue = epilogue;
  const int olast = a.o0 + (D - 1) * a.ostep;
  a.o_lo = a.o0 < olast ? a.o0 : olast;
  const int span = (D - 1) * (disp_step < 0 ? -disp_step : disp_step);
  const int ntm_need = (16 + span + 15) / 16;
  int ntm;
  if (ntm_need <= 3) ntm = 3;
  else if (ntm_need <= 5) ntm = 5;
  else if (ntm_need <= 9) ntm = 9;
  else if (ntm_need <= 17) ntm = 17;
  else return fail(CCS_ERR_UNSUPPORTED, "ccs_correlation_volume: disparity span %d exceeds 256", span);
  // S row pitch: 16*ntm + 3 floats, so that band rows r start (sp+1) r = 20 r
  // (mod 32) words apart and the softmax pass's four-word groups of eight rows
  // fall on distinct banks (with a multiple of 16 the rows met every 2).
  a.sp = 16 * ntm + 3;
  const size_t esz = precision == CCS_BF16 ? 2 : 4;
  a.pitch = precision == CCS_BF16 ? a.Cs : a.Cs + 1;
  // Waves per workgroup: the most resident waves per CU (160 KB of LDS), ties
  // to the wider tile (its f1 window is shared by more pixels).
  int NW = 0, best = 0;
  for (int nw = 8; nw >= 1; nw >>= 1) {
    a.win = 16 * (nw - 1) + 16 * ntm;
    const size_t bytes = corr_lds_bytes(a, nw, esz);
    const int resident = nw * (int)((160 * 1024) / bytes);
    if (resident > best) best = resident, NW = nw;
  }
  CCS_CHECK(NW >= 1, CCS_ERR_UNSUPPORTED, "ccs_correlation_volume: tile does not fit in LDS");
  a.win = 16 * (NW - 1) + 16 * ntm;
  hipStream_t st = h->stream;
  if (precision == CCS_BF16) {
    switch (ntm) {
      case 3: return launch_volume<bf16, 3>(a, NW, st);
      case 5: return launch_volume<bf16, 5>(a, NW, st);
      case 9: return launch_volume<bf16, 9>(a, NW, st);
      default: return launch_volume<bf16, 17>(a, NW, st);
    }
  }
  switch (ntm) {
    case 3: return launch_volume<float, 3>(a, NW, st);
    case 5: return launch_volume<float, 5>(a, NW, st);
    case 9: return launch_volume<float, 9>(a, NW, st);
    default: return launch_volume<float, 17>(a, NW, st);
  }
}

// Reads the reference's float disparity list to the host and checks that every
// entry is an integer (the interpolating kernels of the reference are not offered).
static int fetch_disparities(ccs_handle_t h, const float* d_disp, int n, std::vector<int>& out) {
  std::vector<float> hd(n);
  CCS_HIP(hipMemcpyAsync(hd.data(), d_disp, sizeof(float) * n, hipMemcpyDeviceToHost, h->stream));
  CCS_HIP(hipStreamSynchronize(h->stream));
  out.resize(n);
  for (int i = 0; i < n; ++i) {
    const float r = std::round(hd[i]);
    CCS_CHECK(std::fabs(hd[i] - r) <= 1e-3f, CCS_ERR_UNSUPPORTED,
              "ccs_correlation: fractional disparity %g (interpolating correlation is not provided)", hd[i]);
    out[i] = (int)r;
  }
  return CCS_OK;
}

extern "C" int ccs_correlation_forward(ccs_handle_t h, const float* input0, const float* input1,
                                       const float* disparities, float* output, int in, int ic, int ih, int iw,
                                       int len_disps, int memory_layout, int rect_corr) {
  CCS_CHECK(h && input0 && input1 && disparities && output, CCS_ERR_ARG, "ccs_correlation_forward: NULL argument");
  CCS_CHECK(in >= 1 && ic >= 1 && ih >= 1 && iw >= 1 && len_disps >= 1, CCS_ERR_ARG,
            "ccs_correlation_forward: empty shape");
  CCS_CHECK(memory_layout == CCS_LAYOUT_NHWC || memory_layout == CCS_LAYOUT_NCHW, CCS_ERR_ARG,
            "ccs_correlation_forward: bad memory layout %d", memory_layout);
  CCS_CHECK(rect_corr >= 0, CCS_ERR_ARG, "ccs_correlation_forward: negative rect_corr");
  std::vector<int> disp;
  int rc = fetch_disparities(h, disparities, len_disps, disp);
  if (rc) return rc;
  if ((rc = h->misc.reserve(sizeof(int) * len_disps, h->stream))) return rc;
  int* dd = (int*)h->misc.ptr;
  CCS_HIP(hipMemcpyAsync(dd, disp.data(), sizeof(int) * len_disps, hipMemcpyHostToDevice, h->stream));
  const long total = (long)in * ih * iw * len_disps;
  k_corr_generic<<<ceil_div(total, 256), 256, 0, h->stream>>>(input0, input1, dd, output, in, ic, ih, iw, len_disps,
                                                             memory_layout == CCS_LAYOUT_NHWC, rect_corr);
  CCS_LAUNCH_CHECK();
  CCS_HIP(hipStreamSynchronize(h->stream));
  return CCS_OK;
}

extern "C" int ccs_correlation_backward(ccs_handle_t h, float* out_grad0, float* out_grad1, const float* in_grad,
                                        const float* im0, const float* im1, const float* disparities, int n, int c,
                                        int ih, int iw, int num_disps, int memory_layout) {
  CCS_CHECK(h && out_grad0 && out_grad1 && in_grad && im0 && im1 && disparities, CCS_ERR_ARG,
            "ccs_correlation_backward: NULL argument");
  CCS_CHECK(n >= 1 && c >= 1 && ih >= 1 && iw >= 1 && num_disps >= 1, CCS_ERR_ARG,
            "ccs_correlation_backward: empty shape");
  CCS_CHECK(memory_layout == CCS_LAYOUT_NHWC || memory_layout == CCS_LAYOUT_NCHW, CCS_ERR_ARG,
            "ccs_correlation_backward: bad memory layout");
  std::vector<int> disp;
  int rc = fetch_disparities(h, disparities, num_disps, disp);
  if (rc) return rc;
  if ((rc = h->misc.reserve(sizeof(int) * num_disps, h->stream))) return rc;
  int* dd = (int*)h->misc.ptr;
  CCS_HIP(hipMemcpyAsync(dd, disp.data(), sizeof(int) * num_disps, hipMemcpyHostToDevice, h->stream));
  const long total = (long)n * c * ih * iw;
  k_corr_bwd_generic<<<ceil_div(total, 256), 256, 0, h->stream>>>(out_grad0, out_grad1, in_grad, im0, im1, dd, n, c,
                                                                 ih, iw, num_disps,
                                                                 memory_layout == CCS_LAYOUT_NHWC);
  CCS_LAUNCH_CHECK();
  CCS_HIP(hipStreamSynchronize(h->stream));
  return CCS_OK;
}

extern "C" int ccs_correlation_volume_backward(ccs_handle_t h, const void* f0, const void* f1, int precision, int H,
                                               int W, int C, int disp_min, int disp_step, int D, const float* G,
                                               void* df0, void* df1) {
  CCS_CHECK(h && f0 && f1 && G && df0 && df1, CCS_ERR_ARG, "ccs_correlation_volume_backward: NULL argument");
  CCS_CHECK(H >= 1 && W >= 1 && C >= 1 && D >= 1, CCS_ERR_ARG, "ccs_correlation_volume_backward: empty shape");
  CCS_CHECK(precision == CCS_FP32 || precision == CCS_BF16, CCS_ERR_ARG, "ccs_correlation_volume_backward: precision");
  const int Cs = feature_stride(C);
  const long total = (long)H * W * Cs;
  const int o0 = -disp_min, ostep = -disp_step;
  if (Cs <= 112 && ostep != 0) {
    // MFMA form: two launches (df0 from f1, df1 from f0 with G gathered along the band)
    const int astep = ostep < 0 ? -ostep : ostep;
    const int KC = std::max(1, kBwdKC / astep);
    const int span_max = (KC - 1) * astep;
    const int ct = (Cs + 15) / 16;
    const int CT = ct <= 1 ? 1 : ct <= 2 ? 2 : ct <= 4 ? 4 : 7;  // the instantiated channel tiles
    const size_t lds = sizeof(float) * ((size_t)kBwdTW * KC + (size_t)(kBwdTW + span_max) * 16 * CT) +
                       sizeof(int) * (kBwdTW + span_max + 16);
    const dim3 grid(ceil_div(W, kBwdTW), H);
    auto launch = [&](auto tag, const void* fs, int e0, int es, int gather, void* o) {
      using T = decltype(tag);
      auto go = [&](auto ct) {
        constexpr int NT = decltype(ct)::value;
        auto fn = k_corr_bwd_mfma<T, NT>;
        if (lds > 64 * 1024) (void)hipFuncSetAttribute((const void*)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        fn<<<grid, 256, lds, h->stream>>>((const T*)fs, G, H, W, Cs, D, e0, es, gather, o0, ostep, (T*)o);
      };
      if (CT == 1) go(std::integral_constant<int, 1>{});
      else if (CT == 2) go(std::integral_constant<int, 2>{});
      else if (CT == 4) go(std::integral_constant<int, 4>{});
      else go(std::integral_constant<int, 7>{});
    };
    if (precision == CCS_BF16) {
      launch(bf16{}, f1, o0, ostep, 0, df0);
      launch(bf16{}, f0, -o0, -ostep, 1, df1);
    } else {
      launch(float{}, f1, o0, ostep, 0, df0);
      launch(float{}, f0, -o0, -ostep, 1, df1);
    }
    CCS_LAUNCH_CHECK();
    return CCS_OK;
  }
  if (precision == CCS_BF16)  // VALU form for channel counts beyond the MFMA tiles
    k_corr_volume_bwd<bf16><<<ceil_div(total, 256), 256, 0, h->stream>>>((const bf16*)f0, (const bf16*)f1, H, W, C,
                                                                        Cs, o0, ostep, D, G, (bf16*)df0,
                                                                        (bf16*)df1);
  else
    k_corr_volume_bwd<float><<<ceil_div(total, 256), 256, 0, h->stream>>>((const float*)f0, (const float*)f1, H, W,
                                                                         C, Cs, o0, ostep, D, G, (float*)df0,
                                                                         (float*)df1);
  CCS_LAUNCH_CHECK();
  return CCS_OK;
}

extern "C" int ccs_softmax_xent(ccs_handle_t h, const float* S, int HW, int D, const int* target, double* loss,
                                float* dS) {
  CCS_CHECK(h && S && target && loss && dS && HW >= 1 && D >= 1, CCS_ERR_ARG, "ccs_softmax_xent: bad argument");
  // One wave per pixel at a time: enough blocks for every wave slot of the chip.
  const int nb = (int)std::min<long>(kXentBlocks, ((long)HW + 3) / 4);
  int rc;
  if ((rc = h->misc2.reserve(sizeof(double) * nb, h->stream))) return rc;
  double* part = (double*)h->misc2.ptr;
  k_softmax_xent<<<nb, 256, 0, h->stream>>>(S, HW, D, target, dS, part);
  k_sum_f64_corr<<<1, 256, 0, h->stream>>>(part, nb, loss);
  CCS_LAUNCH_CHECK();
  return CCS_OK;
}
