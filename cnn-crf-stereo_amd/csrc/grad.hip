// Adjoint of the fused convolution + bias + tanh layer (float32 feature maps).
// The reference has no convolution backward; this follows the forward
// definition y = tanh(b + sum_{c,i,j} x[h+i][w+j][c] W[k][c][KH-1-i][KW-1-j]):
//   dz = dy * (1 - y^2)                (y recomputed from x, not stored)
//   db[k] = sum_p dz[p][k]
//   dW[k][c][KH-1-i][KW-1-j] = sum_p dz[p][k] x[p + (i,j)][c]
//   dx = true convolution of dz zero-padded by (KH-1, KW-1) with
//        Wt[c][k][a][b] = W[k][c][KH-1-a][KW-1-b]
// Reductions over pixels are split into blocks and summed deterministically.
#include "common.hpp"

namespace ccs {

__global__ void k_dz(const float* __restrict__ y, const float* __restrict__ dy, long n, int act,
                     float* __restrict__ dz) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float yv = y[i];
  dz[i] = act ? dy[i] * (1.f - yv * yv) : dy[i];
}

// Partial bias gradient: block (b, j) sums pixels [b*chunk, (b+1)*chunk) of
// channels j*256 .. j*256+255.
__global__ void k_db_partial(const float* __restrict__ dz, long P, int Cs, int chunk, float* __restrict__ part) {
  const int c = blockIdx.y * blockDim.x + threadIdx.x;
  if (c >= Cs) return;
  const long p0 = (long)blockIdx.x * chunk;
  const long p1 = p0 + chunk < P ? p0 + chunk : P;
  float s = 0.f;
  for (long p = p0; p < p1; ++p) s += dz[p * Cs + c];
  part[(long)blockIdx.x * Cs + c] = s;
}

__global__ void k_reduce_rows(const float* __restrict__ part, int nb, int n, float* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float s = 0.f;
  for (int b = 0; b < nb; ++b) s += part[(long)b * n + i];
  out[i] = s;
}

// Partial weight gradient.  grid = (ceil(K/16), ceil(C/16) * taps, nb); block
// 16 x 16 threads, thread (tk, tc) owns dW for (k0+tk, c0+tc) at one tap and
// sums its block's pixel range.  Pixels go through LDS 64 at a time.
__global__ __launch_bounds__(256) void k_dw_partial(const float* __restrict__ x, int Win, int Csx,
                                                    const float* __restrict__ dz, int Ho, int Wo, int Csk, int K,
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
        if (k0 + e < K) zv = dz[p * Csk + k0 + e];
        if (c0 + e < C) xv = x[((long)(ho + i) * Win + wo + j) * Csx + c0 + e];
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

__global__ void k_pad_dz(const float* __restrict__ dz, int Ho, int Wo, int Cs, int py, int px,
                         float* __restrict__ out) {
  const int Hp = Ho + 2 * py, Wp = Wo + 2 * px;
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)Hp * Wp * Cs) return;
  const int c = (int)(idx % Cs);
  const long p = idx / Cs;
  const int yy = (int)(p / Wp) - py, xx = (int)(p % Wp) - px;
  out[idx] = (yy >= 0 && yy < Ho && xx >= 0 && xx < Wo) ? dz[((long)yy * Wo + xx) * Cs + c] : 0.f;
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

extern "C" int ccs_conv2d_backward(ccs_handle_t h, const float* x, int H, int W, int C, const float* w,
                                   const float* b, int K, int KH, int KW, int act, const float* dy, float* dx,
                                   float* dw, float* db) {
  CCS_CHECK(h && x && w && dy && dw, CCS_ERR_ARG, "ccs_conv2d_backward: NULL argument");
  CCS_CHECK(C >= 1 && K >= 1 && KH >= 1 && KW >= 1 && H >= KH && W >= KW, CCS_ERR_SHAPE,
            "ccs_conv2d_backward: bad shape");
  const int Ho = H - KH + 1, Wo = W - KW + 1;
  const int Csx = feature_stride(C), Csk = feature_stride(K);
  const long P = (long)Ho * Wo;
  const long nz = P * Csk;
  const int nb = 64;
  const long chunk = (P + nb - 1) / nb;
  const size_t wcount = (size_t)K * C * KH * KW;
  const int Hp = Ho + 2 * (KH - 1), Wp = Wo + 2 * (KW - 1);
  // scratch: y/dz (nz) | partials (nb * max(wcount, Csk)) | padded dz | Wt
  const size_t part_n = (size_t)nb * (wcount > (size_t)Csk ? wcount : (size_t)Csk);
  const size_t pad_n = dx ? (size_t)Hp * Wp * Csk : 0;
  const size_t need = sizeof(float) * (nz + part_n + pad_n + wcount);
  int rc = h->misc2.reserve(need, h->stream);
  if (rc) return rc;
  float* z = (float*)h->misc2.ptr;
  float* part = z + nz;
  float* zpad = part + part_n;
  float* wt = zpad + pad_n;

  if ((rc = ccs_conv2d_forward(h, x, CCS_LAYOUT_NHWC, H, W, C, w, b, K, KH, KW, act, CCS_FP32, z))) return rc;
  k_dz<<<ceil_div(nz, 256), 256, 0, h->stream>>>(z, dy, nz, act, z);
  if (db) {
    k_db_partial<<<dim3(nb, ceil_div(Csk, 256)), 256, 0, h->stream>>>(z, P, Csk, (int)chunk, part);
    k_reduce_rows<<<ceil_div(Csk, 256), 256, 0, h->stream>>>(part, nb, Csk, part + (size_t)nb * Csk);
    CCS_HIP(hipMemcpyAsync(db, part + (size_t)nb * Csk, sizeof(float) * K, hipMemcpyDeviceToDevice, h->stream));
  }
  {
    dim3 grid(ceil_div(K, 16), ceil_div(C, 16) * KH * KW, nb);
    k_dw_partial<<<grid, 256, 0, h->stream>>>(x, W, Csx, z, Ho, Wo, Csk, K, C, KH, KW, chunk, part);
    k_reduce_rows<<<ceil_div((long)wcount, 256), 256, 0, h->stream>>>(part, nb, (int)wcount, dw);
  }
  if (dx) {
    k_pad_dz<<<ceil_div((long)Hp * Wp * Csk, 256), 256, 0, h->stream>>>(z, Ho, Wo, Csk, KH - 1, KW - 1, zpad);
    k_transpose_flip<<<ceil_div((long)wcount, 256), 256, 0, h->stream>>>(w, K, C, KH, KW, wt);
    CCS_LAUNCH_CHECK();
    // dx has C output channels from K input channels.
    if ((rc = ccs_conv2d_forward(h, zpad, CCS_LAYOUT_NHWC, Hp, Wp, K, wt, nullptr, C, KH, KW, 0, CCS_FP32, dx)))
      return rc;
  }
  CCS_LAUNCH_CHECK();
  return CCS_OK;
}
