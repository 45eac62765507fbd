// Adjoint of the fused convolution + bias + tanh layer.  The reference has no
// convolution backward; this follows the forward definition
// y = tanh(b + sum_{c,i,j} x[h+i][w+j][c] W[k][c][KH-1-i][KW-1-j]):
//   dz = dy * (1 - y^2)                (y: the layer's output, kept or recomputed)
//   db[k] = sum_p dz[p][k]
//   dW[k][c][KH-1-i][KW-1-j] = sum_p dz[p][k] x[p + (i,j)][c]
//   dx = true convolution of dz zero-padded by (KH-1, KW-1) with
//        Wt[c][k][a][b] = W[k][c][KH-1-a][KW-1-b]
// dW runs on v_mfma_f32_16x16x4_f32 (k_dw_mfma: per tap, a K x C GEMM reduced
// over pixels), dx on the forward MFMA kernels.  Maps are float32 or bfloat16
// (the BF16 form keeps activations and activation gradients in bf16 and
// accumulates in float32).  Reductions over pixels are split into blocks and
// summed in a fixed order: results do not depend on scheduling.
#include <algorithm>
#include <type_traits>

#include "common.hpp"

namespace ccs {

template <typename T>
__device__ __forceinline__ float ldf(const T* p, long i) {
  return (float)p[i];
}

// dz = dy (1 - y^2) (act) or dy for the K live channels, 0 in the pad
// channels, and the bias-gradient partials: block b takes pixels
// [b chunk, (b+1) chunk), thread c channel c, and writes the sum of its stored
// (rounded) dz values to part[c][b] (channel-major, for k_reduce_cols).
template <typename T>
__global__ void k_dz_db(const T* __restrict__ y, const T* __restrict__ dy, long P, int Cs, int K, int act, int chunk,
                        T* __restrict__ dz, float* __restrict__ part) {
  const long p0 = (long)blockIdx.x * chunk;
  const long p1 = p0 + chunk < P ? p0 + chunk : P;
  for (int c = threadIdx.x; c < Cs; c += blockDim.x) {
    float s = 0.f;
    for (long p = p0; p < p1; ++p) {
      const long i = p * Cs + c;
      const float yv = ldf(y, i), g = ldf(dy, i);
      const T t = (T)(c < K ? (act ? g * (1.f - yv * yv) : g) : 0.f);
      dz[i] = t;
      s += (float)t;
    }
    part[(long)c * gridDim.x + blockIdx.x] = s;
  }
}

// out[c] = sum of the n partials part[c][0..n) (one block per c, a fixed tree).
__global__ void k_reduce_cols(const float* __restrict__ part, int n, float* __restrict__ out) {
  __shared__ float sh[256];
  const float* row = part + (long)blockIdx.x * n;
  float s = 0.f;
  for (int i = threadIdx.x; i < n; i += 256) s += row[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = sh[0];
}

__global__ void k_reduce_rows(const float* __restrict__ part, int nb, int n, float* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float s = 0.f;
  for (int b = 0; b < nb; ++b) s += part[(long)b * n + i];
  out[i] = s;
}

// Weight gradient of one tap on MFMA.  A 1-D grid of taps x chunks blocks,
// one wave per 16-channel tile of K (KT waves).  Blocks b and b + 8 share an
// XCD (and its L2), so a chunk's taps are given to consecutive blocks of one
// XCD: they read the same dz and overlapping inputs at about the same time.
// The pixels are walked in segments of 64 output pixels along a row; a
// segment's dz (64 x Csk) and tap-shifted input (64 x Csx) go to LDS as
// float32, channel pitch 16 KT / 16 CT (the pad channels stay zero), while the
// next segment travels in registers.  Per 4 pixels each wave issues CT
// v_mfma_f32_16x16x4_f32: A[m][r] = dz[p_r][k0+m], B[r][n] = x[p_r + tap][16t+n].
// The block's K x C sums go to part[chunk][tap].  Register budget: four waves
// per SIMD, so two 7-wave blocks share a CU (LDS: 2 x 57 KB for 7 x 7 tiles).
template <typename TX, typename T, int KT, int CT>
__global__ __launch_bounds__(64 * KT, 4) void k_dw_mfma(const TX* __restrict__ x, int Win, int Csx,
                                                        const T* __restrict__ dz, int Ho, int Wo, int Csk, int KW,
                                                        int taps, int nch, int segs_per_row, long nseg, long spc,
                                                        float* __restrict__ part) {
  constexpr int KP = 16 * KT, CP = 16 * CT, NTH = 64 * KT, SEG = 64;
  constexpr int VE = 16 / (int)sizeof(T), VX = 16 / (int)sizeof(TX);  // elements per 16-byte vector
  constexpr int VMIN = VE < VX ? VE : VX;
  constexpr int NVMAX = (SEG * (KP + CP) / VMIN + NTH - 1) / NTH;
  __shared__ __align__(16) float zs[SEG * KP];
  __shared__ __align__(16) float xs[SEG * CP];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  // (tap, chunk) of this block: taps of one chunk on consecutive blocks of one XCD.
  const int b = blockIdx.x, xcd = b & 7, local = b >> 3;
  const int tap = local % taps, chunk = (local / taps) * 8 + xcd;
  if (chunk >= nch) return;  // the grid is rounded up to whole groups of 8 chunks
  const int ti = tap / KW, tj = tap % KW;
  const long s0 = (long)chunk * spc;
  const long s1 = s0 + spc < nseg ? s0 + spc : nseg;
  for (int i = tid; i < SEG * KP; i += NTH) zs[i] = 0.f;
  for (int i = tid; i < SEG * CP; i += NTH) xs[i] = 0.f;
  const int zvp = Csk / VE, xvp = Csx / VX;  // vectors per pixel
  const int nzv = SEG * zvp, nv = nzv + SEG * xvp;
  // Per vector slot of this thread (independent of the segment): the pixel in
  // the segment, the element offset within the pixel, the LDS offset and the
  // operand (0: dz, 1: x, -1: none).
  int vpx[NVMAX], voff[NVMAX], vlds[NVMAX], vkind[NVMAX];
#pragma unroll
  for (int i = 0; i < NVMAX; ++i) {
    const int v = tid + i * NTH;
    if (v < nzv) {
      vpx[i] = v / zvp;
      voff[i] = (v % zvp) * VE;
      vlds[i] = vpx[i] * KP + voff[i];
      vkind[i] = 0;
    } else if (v < nv) {
      vpx[i] = (v - nzv) / xvp;
      voff[i] = ((v - nzv) % xvp) * VX;
      vlds[i] = vpx[i] * CP + voff[i];
      vkind[i] = 1;
    } else {
      vpx[i] = voff[i] = vlds[i] = 0;
      vkind[i] = -1;
    }
  }
  u32x4 reg[NVMAX];
  auto load = [&](long s) {
    const int ho = (int)(s / segs_per_row), wo0 = (int)(s % segs_per_row) * SEG;
    const T* zrow = dz + ((long)ho * Wo + wo0) * Csk;
    const TX* xrow = x + ((long)(ho + ti) * Win + wo0 + tj) * Csx;
#pragma unroll
    for (int i = 0; i < NVMAX; ++i) {
      const u32x4* src = vkind[i] == 0 ? reinterpret_cast<const u32x4*>(zrow + (long)vpx[i] * Csk + voff[i])
                                       : reinterpret_cast<const u32x4*>(xrow + (long)vpx[i] * Csx + voff[i]);
      u32x4 r = {0u, 0u, 0u, 0u};
      if (vkind[i] >= 0 && wo0 + vpx[i] < Wo) r = *src;
      reg[i] = r;
    }
  };
  // 16 bytes of float32 (4 values) or bfloat16 (8 values, two per dword, low
  // half first) as float32 into LDS.
  auto put = [](float* d, const u32x4& r, int ve) {
    if (ve == 4) {
      *reinterpret_cast<u32x4*>(d) = r;
    } else {
      const u32x4 lo = {r[0] << 16, r[0] & 0xffff0000u, r[1] << 16, r[1] & 0xffff0000u};
      const u32x4 hi = {r[2] << 16, r[2] & 0xffff0000u, r[3] << 16, r[3] & 0xffff0000u};
      reinterpret_cast<u32x4*>(d)[0] = lo;
      reinterpret_cast<u32x4*>(d)[1] = hi;
    }
  };
  auto store = [&]() {
#pragma unroll
    for (int i = 0; i < NVMAX; ++i) {
      if (vkind[i] == 0) put(zs + vlds[i], reg[i], VE);
      else if (vkind[i] == 1) put(xs + vlds[i], reg[i], VX);
    }
  };
  f32x4 acc[CT];
#pragma unroll
  for (int t = 0; t < CT; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int lrow = lane & 15, lk = lane >> 4, k0 = 16 * wave;
  if (s0 < s1) load(s0);
  for (long s = s0; s < s1; ++s) {
    __syncthreads();  // every wave is done with the previous segment
    store();
    __syncthreads();
    if (s + 1 < s1) load(s + 1);  // in flight during the MFMAs
#pragma unroll 4
    for (int st = 0; st < SEG / 4; ++st) {
      const int p = 4 * st + lk;
      const float a = zs[p * KP + k0 + lrow];
#pragma unroll
      for (int t = 0; t < CT; ++t)
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, xs[p * CP + 16 * t + lrow], acc[t], 0, 0, 0);
    }
  }
  float* out = part + ((long)chunk * taps + tap) * KP * CP;
#pragma unroll
  for (int t = 0; t < CT; ++t)
#pragma unroll
    for (int e = 0; e < 4; ++e) out[(k0 + 4 * lk + e) * CP + 16 * t + lrow] = acc[t][e];
}

// Weight gradient of one tap row on bf16 MFMA (bf16 input and dz): the block
// handles the KW taps (i, 0 .. KW-1) of tap row i for a chunk of 64-pixel
// segments, one wave per 16-channel tile of K (KT waves), KW x CT accumulator
// tiles per wave.  A segment's dz and its input row go to LDS transposed
// (channel-major, pixel-minor bf16), the input once per tap shift j (so every
// fragment is a 16-byte aligned read): per 32 pixels a wave issues KW x CT
// v_mfma_f32_16x16x32_bf16 with A[m][p] = dz[p][k0+m], B[p][n] = x[p + j][16t+n].
// Two LDS buffers alternate: while the MFMAs run on one segment, the next one's
// vectors (loaded a segment earlier) are written into the other.  The K x C
// sums of tap (i, j) go to part[chunk][tap], the layout of k_dw_mfma.
template <int KT, int CT, int KW>
__global__ __launch_bounds__(64 * KT, 1) void k_dw_bf16(const bf16* __restrict__ x, int Win, int Csx,
                                                        const bf16* __restrict__ dz, int Ho, int Wo, int Csk, int KH,
                                                        int nch, int segs_per_row, long nseg, long spc,
                                                        float* __restrict__ part) {
  constexpr int KP = 16 * KT, CP = 16 * CT, NTH = 64 * KT, SEG = 64, XPX = SEG + KW - 1;
  constexpr int ZB = KP * SEG, XB = KW * CP * SEG;  // bf16 elements of one buffer's dz and x parts
  extern __shared__ __align__(16) unsigned char lds_raw[];
  bf16* lds = reinterpret_cast<bf16*>(lds_raw);  // [buf][ZB + XB]
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int b = blockIdx.x, xcd = b & 7, local = b >> 3;
  const int ti = local % KH, chunk = (local / KH) * 8 + xcd;
  if (chunk >= nch) return;  // the grid is rounded up to whole groups of 8 chunks
  const long s0 = (long)chunk * spc;
  const long s1 = s0 + spc < nseg ? s0 + spc : nseg;
  for (int i = tid; i < 2 * (ZB + XB) / 2; i += NTH) reinterpret_cast<unsigned*>(lds)[i] = 0u;  // pad channels stay 0
  const int zvp = Csk / 8, xvp = Csx / 8;  // 16-byte vectors per pixel
  const int nzv = SEG * zvp, nv = nzv + XPX * xvp;
  constexpr int NVMAX = (SEG * (KP / 8) + XPX * (CP / 8) + NTH - 1) / NTH;
  // Per vector slot: pixel, channel group, operand (0 dz, 1 x, -1 none).
  // Consecutive slots take consecutive pixels of one channel group, so the
  // transposed 16-bit stores of a wave fill contiguous 128-byte row pieces (no
  // bank conflicts); the group's strided global reads hit the L1 lines that the
  // segment's other groups bring in.
  int vpx[NVMAX], vg[NVMAX], vkind[NVMAX];
#pragma unroll
  for (int i = 0; i < NVMAX; ++i) {
    const int v = tid + i * NTH;
    if (v < nzv) {
      vpx[i] = v % SEG; vg[i] = v / SEG; vkind[i] = 0;
    } else if (v < nv) {
      vpx[i] = (v - nzv) % XPX; vg[i] = (v - nzv) / XPX; vkind[i] = 1;
    } else {
      vpx[i] = vg[i] = 0; vkind[i] = -1;
    }
  }
  u32x4 reg[NVMAX];
  auto load = [&](long s) {
    const int ho = (int)(s / segs_per_row), wo0 = (int)(s % segs_per_row) * SEG;
    const bf16* zrow = dz + ((long)ho * Wo + wo0) * Csk;
    const bf16* xrow = x + ((long)(ho + ti) * Win + wo0) * Csx;
#pragma unroll
    for (int i = 0; i < NVMAX; ++i) {
      u32x4 r = {0u, 0u, 0u, 0u};
      if (vkind[i] == 0) {
        if (wo0 + vpx[i] < Wo) r = *reinterpret_cast<const u32x4*>(zrow + (long)vpx[i] * Csk + vg[i] * 8);
      } else if (vkind[i] == 1) {
        if (wo0 + vpx[i] < Win) r = *reinterpret_cast<const u32x4*>(xrow + (long)vpx[i] * Csx + vg[i] * 8);
      }
      reg[i] = r;
    }
  };
  // Transposed stores: element e of a vector (channel 8 g + e) goes to row
  // 8 g + e at the vector's pixel; an input vector at pixel p' goes to shift
  // copy j at pixel p' - j.
  auto store = [&](int buf) {
    bf16* zb = lds + buf * (ZB + XB);
    bf16* xb = zb + ZB;
#pragma unroll
    for (int i = 0; i < NVMAX; ++i) {
      const unsigned short* e16 = reinterpret_cast<const unsigned short*>(&reg[i]);
      if (vkind[i] == 0) {
        unsigned short* d = reinterpret_cast<unsigned short*>(zb) + (vg[i] * 8) * SEG + vpx[i];
#pragma unroll
        for (int e = 0; e < 8; ++e) d[e * SEG] = e16[e];
      } else if (vkind[i] == 1) {
#pragma unroll
        for (int j = 0; j < KW; ++j) {
          const int p = vpx[i] - j;
          if (p >= 0 && p < SEG) {
            unsigned short* d = reinterpret_cast<unsigned short*>(xb) + (j * CP + vg[i] * 8) * SEG + p;
#pragma unroll
            for (int e = 0; e < 8; ++e) d[e * SEG] = e16[e];
          }
        }
      }
    }
  };
  f32x4 acc[KW][CT];
#pragma unroll
  for (int j = 0; j < KW; ++j)
#pragma unroll
    for (int t = 0; t < CT; ++t) acc[j][t] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int lrow = lane & 15, kg = lane >> 4, k0 = 16 * wave;
  __syncthreads();  // the zero fill is done
  if (s0 < s1) {
    load(s0);
    store(0);
    if (s0 + 1 < s1) load(s0 + 1);
  }
  __syncthreads();
  for (long s = s0; s < s1; ++s) {
    const int buf = (int)((s - s0) & 1);
    const bf16* zb = lds + buf * (ZB + XB);
    const bf16* xb = zb + ZB;
    // the next segment into the other buffer (its previous readers finished
    // before the barrier that ended the last iteration)
    if (s + 1 < s1) {
      store(buf ^ 1);
      if (s + 2 < s1) load(s + 2);
    }
#pragma unroll
    for (int h = 0; h < SEG / 32; ++h) {
      const bf16x8 a = *reinterpret_cast<const bf16x8*>(zb + (k0 + lrow) * SEG + 32 * h + 8 * kg);
#pragma unroll
      for (int j = 0; j < KW; ++j)
#pragma unroll
        for (int t = 0; t < CT; ++t) {
          const bf16x8 bv = *reinterpret_cast<const bf16x8*>(xb + (j * CP + 16 * t + lrow) * SEG + 32 * h + 8 * kg);
          acc[j][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bv, acc[j][t], 0, 0, 0);
        }
    }
    __syncthreads();
  }
#pragma unroll
  for (int j = 0; j < KW; ++j) {
    float* out = part + ((long)chunk * KH * KW + ti * KW + j) * KP * CP;
#pragma unroll
    for (int t = 0; t < CT; ++t)
#pragma unroll
      for (int e = 0; e < 4; ++e) out[(k0 + 4 * kg + e) * CP + 16 * t + lrow] = acc[j][t][e];
  }
}

// dW[k][c][KH-1-i][KW-1-j] = sum over chunks of the (tap = i KW + j) partials.
__global__ void k_dw_reduce(const float* __restrict__ part, int nch, int taps, int KP, int CP, int K, int C, int KH,
                            int KW, float* __restrict__ dw) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)K * C * taps) return;
  const int c = (int)(idx % C);
  const int k = (int)((idx / C) % K);
  const int tap = (int)(idx / ((long)C * K));
  float s = 0.f;
  for (int ch = 0; ch < nch; ++ch) s += part[(((long)ch * taps + tap) * KP + k) * CP + c];
  const int i = tap / KW, j = tap % KW;
  dw[(((long)k * C + c) * KH + (KH - 1 - i)) * KW + (KW - 1 - j)] = s;
}

// Partial weight gradient on the VALU, for channel counts beyond the MFMA
// kernel's tiles.  grid = (ceil(K/16), ceil(C/16) * taps, nb); block 16 x 16
// threads, thread (tk, tc) owns dW for (k0+tk, c0+tc) at one tap and sums its
// block's pixel range.  Pixels go through LDS 64 at a time.
template <typename TX, typename T>
__global__ __launch_bounds__(256) void k_dw_partial(const TX* __restrict__ x, int Win, int Csx,
                                                    const T* __restrict__ dz, int Ho, int Wo, int Csk, int K,
                                                    int C, int KH, int KW, long chunk, float* __restrict__ part) {
  const int tk = threadIdx.x & 15, tc = threadIdx.x >> 4;
  const int k0 = blockIdx.x * 16;
  const int cblocks = (C + 15) / 16;
  const int tap = blockIdx.y / cblocks, c0 = (blockIdx.y % cblocks) * 16;
  const int i = tap / KW, j = tap % KW;
  const long P = (long)Ho * Wo;
  const long p0 = (long)blockIdx.z * chunk;
  const long p1 = p0 + chunk < P ? p0 + chunk : P;
  __shared__ float zs[64][17], xs[64][17];
  float acc = 0.f;
  for (long pb = p0; pb < p1; pb += 64) {
    for (int v = threadIdx.x; v < 64 * 16; v += 256) {
      const int pl = v >> 4, e = v & 15;
      const long p = pb + pl;
      float zv = 0.f, xv = 0.f;
      if (p < p1) {
        const int ho = (int)(p / Wo), wo = (int)(p % Wo);
        if (k0 + e < K) zv = ldf(dz, p * Csk + k0 + e);
        if (c0 + e < C) xv = ldf(x, ((long)(ho + i) * Win + wo + j) * Csx + c0 + e);
      }
      zs[pl][e] = zv;
      xs[pl][e] = xv;
    }
    __syncthreads();
#pragma unroll 8
    for (int pl = 0; pl < 64; ++pl) acc += zs[pl][tk] * xs[pl][tc];
    __syncthreads();
  }
  const int k = k0 + tk, c = c0 + tc;
  if (k < K && c < C)
    part[(long)blockIdx.z * K * C * KH * KW + (((long)k * C + c) * KH + (KH - 1 - i)) * KW + (KW - 1 - j)] = acc;
}

template <typename T>
__global__ void k_pad_dz(const T* __restrict__ dz, int Ho, int Wo, int Cs, int py, int px, T* __restrict__ out) {
  const int Hp = Ho + 2 * py, Wp = Wo + 2 * px;
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)Hp * Wp * Cs) return;
  const int c = (int)(idx % Cs);
  const long p = idx / Cs;
  const int yy = (int)(p / Wp) - py, xx = (int)(p % Wp) - px;
  out[idx] = (yy >= 0 && yy < Ho && xx >= 0 && xx < Wo) ? dz[((long)yy * Wo + xx) * Cs + c] : (T)0.f;
}

// Wt[c][k][a][b] = W[k][c][KH-1-a][KW-1-b]
__global__ void k_transpose_flip(const float* __restrict__ w, int K, int C, int KH, int KW, float* __restrict__ wt) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long total = (long)K * C * KH * KW;
  if (idx >= total) return;
  const int b = (int)(idx % KW);
  const int a = (int)((idx / KW) % KH);
  const int k = (int)((idx / ((long)KW * KH)) % K);
  const int c = (int)(idx / ((long)KW * KH * K));
  wt[idx] = w[(((long)k * C + c) * KH + (KH - 1 - a)) * KW + (KW - 1 - b)];
}

}  // namespace ccs

using namespace ccs;

namespace {

template <typename TX, typename T, int KT, int CT>
void launch_dw_mfma(const TX* x, int W, int Csx, const T* z, int Ho, int Wo, int Csk, int KW, int taps,
                    int segs_row, long nseg, long spc, int nch, float* part, hipStream_t st) {
  const unsigned blocks = (unsigned)(8 * taps * ((nch + 7) / 8));  // whole groups of 8 chunks, one per XCD
  k_dw_mfma<TX, T, KT, CT><<<blocks, 64 * KT, 0, st>>>(x, W, Csx, z, Ho, Wo, Csk, KW, taps, nch, segs_row, nseg, spc,
                                                      part);
}

// Tiles of the MFMA weight gradient: the smallest instantiated count >= n.
int dw_tiles(int n) { return n <= 1 ? 1 : n <= 2 ? 2 : n <= 4 ? 4 : n <= 7 ? 7 : 0; }

template <typename TX, typename T>
bool dw_mfma(int KT, int CT, const TX* x, int W, int Csx, const T* z, int Ho, int Wo, int Csk, int KW, int taps,
             int segs_row, long nseg, long spc, int nch, float* part, hipStream_t st) {
#define CCS_DW(kt, ct)                                                                                      \
  if (KT == kt && CT == ct) {                                                                               \
    launch_dw_mfma<TX, T, kt, ct>(x, W, Csx, z, Ho, Wo, Csk, KW, taps, segs_row, nseg, spc, nch, part, st); \
    return true;                                                                                            \
  }
  CCS_DW(7, 7) CCS_DW(7, 1) CCS_DW(4, 4) CCS_DW(4, 1) CCS_DW(1, 1) CCS_DW(2, 2) CCS_DW(2, 1) CCS_DW(1, 2)
  CCS_DW(7, 4) CCS_DW(4, 7) CCS_DW(7, 2) CCS_DW(2, 7) CCS_DW(4, 2) CCS_DW(2, 4) CCS_DW(1, 4) CCS_DW(1, 7)
#undef CCS_DW
  return false;
}

constexpr int kDwSeg = 64;            // output pixels per segment of k_dw_mfma and k_dw_bf16
constexpr int kDwTargetBlocks = 512;  // two resident blocks per CU of the 256: one round
constexpr int kDwBf16Blocks = 256;    // k_dw_bf16: one 114 KB block per CU

void launch_dw_bf16(const bf16* x, int W, int Csx, const bf16* z, int Ho, int Wo, int Csk, int KH, int KW,
                    int segs_row, long nseg, long spc, int nch, float* part, hipStream_t st) {
  const unsigned blocks = (unsigned)(8 * KH * ((nch + 7) / 8));  // whole groups of 8 chunks, one per XCD
  auto go = [&](auto kwc) {
    constexpr int kw = decltype(kwc)::value;
    constexpr size_t lds = 2 * sizeof(bf16) * (112 * 64 + kw * 112 * 64);
    const auto kern = k_dw_bf16<7, 7, kw>;
    static bool configured = false;
    if (!configured) {
      (void)hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
      configured = true;
    }
    kern<<<blocks, 64 * 7, lds, st>>>(x, W, Csx, z, Ho, Wo, Csk, KH, nch, segs_row, nseg, spc, part);
  };
  if (KW == 3)
    go(std::integral_constant<int, 3>{});
  else
    go(std::integral_constant<int, 2>{});
}

// The backward of one layer from its input x (element type TX) and output y
// (element type T of the precision); dy has y's shape and type, dx x's shape
// and y's type.
template <typename TX, typename T>
int conv_backward(ccs_handle_t h, const TX* x, const T* y, int H, int W, int C, const float* w, int K, int KH,
                  int KW, int act, const T* dy, T* dx, float* dw, float* db, int precision) {
  const int Ho = H - KH + 1, Wo = W - KW + 1;
  const int Csx = feature_stride(C), Csk = feature_stride(K);
  const long P = (long)Ho * Wo;
  const long nz = P * Csk;
  const int taps = KH * KW;
  const size_t wcount = (size_t)K * C * KH * KW;
  const int Hp = Ho + 2 * (KH - 1), Wp = Wo + 2 * (KW - 1);
  const int KT = dw_tiles((K + 15) / 16), CT = dw_tiles((C + 15) / 16);
  const bool mfma = KT && CT && Csk <= 16 * KT && Csx <= 16 * CT;
  const int segs_row = (Wo + kDwSeg - 1) / kDwSeg;
  const long nseg = (long)Ho * segs_row;
  // bf16 input and dz with 7 x 7 channel tiles: the tap-row kernel on bf16 MFMA
  // (one block per tap row and chunk, about one block per CU in all).
  const bool dw_bf16 = std::is_same<TX, bf16>::value && std::is_same<T, bf16>::value && mfma && KT == 7 &&
                       CT == 7 && (KW == 2 || KW == 3);
  const int dw_units = dw_bf16 ? KH : taps;  // blocks per chunk
  int nch = (int)std::min<long>(nseg, std::max(8, (dw_bf16 ? kDwBf16Blocks : kDwTargetBlocks) / dw_units / 8 * 8));
  const long spc = (nseg + nch - 1) / nch;
  nch = (int)((nseg + spc - 1) / spc);
  const int nb = 64;  // pixel blocks of the VALU weight gradient
  const long chunk = (P + nb - 1) / nb;
  const int dz_chunk = 64;  // pixels per block of the bias-gradient partials
  const int nbz = (int)((P + dz_chunk - 1) / dz_chunk);
  size_t part_n = (size_t)nbz * Csk + Csk;
  part_n = std::max(part_n, mfma ? (size_t)nch * taps * 256 * KT * CT : (size_t)nb * wcount);
  const size_t pad_n = dx ? (size_t)Hp * Wp * Csk : 0;
  auto al = [](size_t b) { return (b + 255) / 256 * 256; };
  const size_t need = al(sizeof(T) * nz) + al(sizeof(float) * part_n) + al(sizeof(T) * pad_n) + sizeof(float) * wcount;
  int rc = h->misc2.reserve(need, h->stream);
  if (rc) return rc;
  char* base = (char*)h->misc2.ptr;
  T* z = (T*)base;
  float* part = (float*)(base + al(sizeof(T) * nz));
  T* zpad = (T*)((char*)part + al(sizeof(float) * part_n));
  float* wt = (float*)((char*)zpad + al(sizeof(T) * pad_n));

  k_dz_db<T><<<nbz, Csk <= 256 ? (Csk + 63) / 64 * 64 : 256, 0, h->stream>>>(y, dy, P, Csk, K, act, dz_chunk, z,
                                                                                part);
  if (db) {
    k_reduce_cols<<<Csk, 256, 0, h->stream>>>(part, nbz, part + (size_t)nbz * Csk);
    CCS_HIP(hipMemcpyAsync(db, part + (size_t)nbz * Csk, sizeof(float) * K, hipMemcpyDeviceToDevice, h->stream));
  }
  if (dw_bf16) {
    if constexpr (std::is_same<TX, bf16>::value && std::is_same<T, bf16>::value)
      launch_dw_bf16(x, W, Csx, z, Ho, Wo, Csk, KH, KW, segs_row, nseg, spc, nch, part, h->stream);
    k_dw_reduce<<<ceil_div((long)wcount, 256), 256, 0, h->stream>>>(part, nch, taps, 16 * KT, 16 * CT, K, C, KH, KW,
                                                                  dw);
  } else if (mfma) {
    dw_mfma<TX, T>(KT, CT, x, W, Csx, z, Ho, Wo, Csk, KW, taps, segs_row, nseg, spc, nch, part, h->stream);
    k_dw_reduce<<<ceil_div((long)wcount, 256), 256, 0, h->stream>>>(part, nch, taps, 16 * KT, 16 * CT, K, C, KH, KW,
                                                                  dw);
  } else {
    dim3 grid(ceil_div(K, 16), ceil_div(C, 16) * KH * KW, nb);
    k_dw_partial<TX, T><<<grid, 256, 0, h->stream>>>(x, W, Csx, z, Ho, Wo, Csk, K, C, KH, KW, chunk, part);
    k_reduce_rows<<<ceil_div((long)wcount, 256), 256, 0, h->stream>>>(part, nb, (int)wcount, dw);
  }
  if (dx) {
    k_pad_dz<T><<<ceil_div((long)Hp * Wp * Csk, 256), 256, 0, h->stream>>>(z, Ho, Wo, Csk, KH - 1, KW - 1, zpad);
    k_transpose_flip<<<ceil_div((long)wcount, 256), 256, 0, h->stream>>>(w, K, C, KH, KW, wt);
    CCS_LAUNCH_CHECK();
    // dx has C output channels from K input channels.
    if ((rc = ccs_conv2d_forward(h, zpad, CCS_LAYOUT_NHWC, Hp, Wp, K, wt, nullptr, C, KH, KW, 0, precision, dx)))
      return rc;
  }
  CCS_LAUNCH_CHECK();
  return CCS_OK;
}

int check_backward(const void* x, const float* w, const void* dy, float* dw, int H, int W, int C, int K, int KH,
                   int KW, const char* who) {
  CCS_CHECK(x && w && dy && dw, CCS_ERR_ARG, "%s: NULL argument", who);
  CCS_CHECK(C >= 1 && K >= 1 && KH >= 1 && KW >= 1 && H >= KH && W >= KW, CCS_ERR_SHAPE, "%s: bad shape", who);
  CCS_CHECK(C <= 4096 && K <= 4096, CCS_ERR_UNSUPPORTED, "%s: too many channels", who);
  return CCS_OK;
}

}  // namespace

extern "C" int ccs_conv2d_backward_stored(ccs_handle_t h, const void* x, int x_precision, const void* y,
                                          int precision, int H, int W, int C, const float* w, int K, int KH, int KW,
                                          int act, const void* dy, void* dx, float* dw, float* db) {
  CCS_CHECK(h && y, CCS_ERR_ARG, "ccs_conv2d_backward_stored: NULL argument");
  int rc = check_backward(x, w, dy, dw, H, W, C, K, KH, KW, "ccs_conv2d_backward_stored");
  if (rc) return rc;
  CCS_CHECK((precision == CCS_FP32 || precision == CCS_BF16) && (x_precision == CCS_FP32 || x_precision == CCS_BF16),
            CCS_ERR_ARG, "ccs_conv2d_backward_stored: bad precision");
  CCS_CHECK(!(x_precision == CCS_BF16 && precision == CCS_FP32), CCS_ERR_UNSUPPORTED,
            "ccs_conv2d_backward_stored: a bfloat16 input under a float32 layer");
  if (precision == CCS_BF16 && x_precision == CCS_BF16)
    return conv_backward<bf16, bf16>(h, (const bf16*)x, (const bf16*)y, H, W, C, w, K, KH, KW, act, (const bf16*)dy,
                                     (bf16*)dx, dw, db, precision);
  if (precision == CCS_BF16)  // float32 input (the image) under a bfloat16 layer
    return conv_backward<float, bf16>(h, (const float*)x, (const bf16*)y, H, W, C, w, K, KH, KW, act,
                                      (const bf16*)dy, (bf16*)dx, dw, db, precision);
  return conv_backward<float, float>(h, (const float*)x, (const float*)y, H, W, C, w, K, KH, KW, act,
                                     (const float*)dy, (float*)dx, dw, db, precision);
}

extern "C" int ccs_conv2d_backward(ccs_handle_t h, const float* x, int H, int W, int C, const float* w,
                                   const float* b, int K, int KH, int KW, int act, const float* dy, float* dx,
                                   float* dw, float* db) {
  CCS_CHECK(h, CCS_ERR_ARG, "ccs_conv2d_backward: NULL argument");
  int rc = check_backward(x, w, dy, dw, H, W, C, K, KH, KW, "ccs_conv2d_backward");
  if (rc) return rc;
  // The layer's output is recomputed from x into the handle's conv_y scratch,
  // then the stored-output path runs.
  const int Ho = H - KH + 1, Wo = W - KW + 1;
  if ((rc = h->conv_y.reserve(sizeof(float) * (size_t)Ho * Wo * feature_stride(K), h->stream))) return rc;
  float* y = (float*)h->conv_y.ptr;
  if ((rc = ccs_conv2d_forward(h, x, CCS_LAYOUT_NHWC, H, W, C, w, b, K, KH, KW, act, CCS_FP32, y))) return rc;
  return conv_backward<float, float>(h, x, y, H, W, C, w, K, KH, KW, act, dy, dx, dw, db, CCS_FP32);
}
