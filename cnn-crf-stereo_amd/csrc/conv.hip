// Feature-CNN layers: fused valid convolution (kernel flipped, CUDNN_CONVOLUTION
// semantics of stereo/operators/convolution.cpp:108,149) + bias + tanh.
//
// Feature maps are channel-last, H x W x Cs with Cs = C rounded up to 8, so the
// im2col row of an output pixel is KH runs of KW * Cs contiguous elements and
// every 8-element group (16 bytes in bf16) stays inside one tap.  The
// reduction index runs over (tap, c) with tap = i * KW + j, and the packed
// weight holds W[n][c][KH-1-i][KW-1-j] there (the flip of true convolution);
// the bf16 packing orders it by channel stage (see stage_count below).
//
// k_conv_bf16_st: the 100-channel layers (3x3 and 2x2) on
//   v_mfma_f32_16x16x32_bf16: a workgroup of 4 waves computes a 4-row x
//   64-pixel output tile for all 112 (padded) output channels, the input halo
//   tile staged in LDS one channel half at a time, the weights streaming
//   through a ring of chunk slots filled by direct global->LDS loads.
//   Epilogue: + bias (the accumulators' start value), tanh, bf16, stored from
//   the accumulators.
// k_conv_bf16: the same GEMM for the other channel counts (64 and 8): 4-row x
//   TW-pixel tiles, the whole input tile in LDS, weights in chunk groups
//   staged through registers, output staged through LDS.
// k_conv_f32: the same on v_mfma_f32_16x16x4_f32 (2-row tile, float32 maps).
// k_conv_generic: VALU direct convolution for the 1- and 3-channel input layer
//   and any shape the MFMA kernels do not cover.
#include <cmath>

#include <mutex>
#include <type_traits>

#include "common.hpp"

namespace ccs {

// ---------------------------------------------------------------------------
// Weight packing

// bf16 layout: [q][n][32] -- q = chunk of 32 reduction indices, n < 16*NT.  The
// four 8-element groups of row n are stored XOR-swizzled by (n >> 1) & 3, so the
// 16-byte LDS reads of eight consecutive rows land in eight distinct bank
// groups without padding the rows.
constexpr int RPB = 32;
__host__ __device__ __forceinline__ int bswz(int n, int g) { return g ^ ((n >> 1) & 3); }
// f32 layout: [q][kk < 4][n] (n contiguous).

__device__ __forceinline__ float flipped_weight(const float* __restrict__ w, int n, int c, int tap, int K, int C,
                                                int KH, int KW) {
  const int taps = KH * KW;
  if (n >= K || c >= C || tap >= taps) return 0.f;
  const int i = tap / KW, j = tap % KW;
  return w[(((long)n * C + c) * KH + (KH - 1 - i)) * KW + (KW - 1 - j)];
}

// Channel stages of the bf16 layout.  The 13 eight-channel groups of the
// 100-channel maps are split 7 + 6 (a workgroup stages half of the input
// channels at a time); other channel counts form one stage.  Chunk q of stage s
// holds the stage's groups kg = 4 (q - q0(s)) + 0..3 in tap-major order
// (kg = tap * ng(s) + g, channel group cg0(s) + g); groups past taps * ng(s) are
// zero padding.  With one stage this is the plain order k = tap * Cs + c.
__host__ __device__ constexpr int stage_count(int CG) { return CG == 13 ? 2 : 1; }
__host__ __device__ constexpr int stage_ng(int CG, int s) { return CG == 13 ? (s == 0 ? 7 : 6) : CG; }
__host__ __device__ constexpr int stage_cg0(int CG, int s) { return CG == 13 && s == 1 ? 7 : 0; }
__host__ __device__ constexpr int stage_nq(int CG, int taps, int s) { return (taps * stage_ng(CG, s) + 3) / 4; }
__host__ __device__ constexpr int packed_chunks(int CG, int taps) {
  return stage_nq(CG, taps, 0) + (stage_count(CG) > 1 ? stage_nq(CG, taps, 1) : 0);
}

__global__ void k_pack_bf16(const float* __restrict__ w, int K, int C, int Cs, int KH, int KW, int NQ, int NP,
                            bf16* __restrict__ out) {
  const long total = (long)NQ * NP * RPB;
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= total) return;
  const int kk = (int)(idx % RPB);
  const int n = (int)((idx / RPB) % NP);
  const int q = (int)(idx / ((long)RPB * NP));
  const int CG = Cs / 8, taps = KH * KW;
  const int st = stage_count(CG) > 1 && q >= stage_nq(CG, taps, 0) ? 1 : 0;
  const int ng = stage_ng(CG, st);
  const int kg = 4 * (q - (st ? stage_nq(CG, taps, 0) : 0)) + bswz(n, kk / 8);  // the swizzle is an involution
  const int tap = kg < taps * ng ? kg / ng : taps;  // taps: a padding group (zero)
  const int c = (stage_cg0(CG, st) + kg % ng) * 8 + kk % 8;
  out[idx] = f32_to_bf16(flipped_weight(w, n, c, tap, K, C, KH, KW));
}

__global__ void k_pack_f32(const float* __restrict__ w, int K, int C, int Cs, int KH, int KW, int NQ, int NP,
                           float* __restrict__ out) {
  const long total = (long)NQ * 4 * NP;
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= total) return;
  const int n = (int)(idx % NP);
  const int kk = (int)((idx / NP) % 4);
  const int q = (int)(idx / ((long)NP * 4));
  const int k = q * 4 + kk;
  out[idx] = flipped_weight(w, n, k % Cs, k / Cs, K, C, KH, KW);
}

// ---------------------------------------------------------------------------
// BF16 MFMA implicit GEMM

constexpr int kTH = 4;    // output rows per tile (one per wave)
#ifndef CCS_CONV_SCHED
#define CCS_CONV_SCHED 1
#endif
constexpr int kLdsBytes = 160 * 1024;

// TW output pixels per row; the input tile plus one weight group fit in
// LDS_BUDGET bytes so that 160 KB / LDS_BUDGET workgroups share a CU.
template <int KH, int KW, int CG, int NT, int TW, int LDS_BUDGET>
struct Bf16Cfg {
  static constexpr int Cs = CG * 8;
  static constexpr int MT = TW / 16;              // M-tiles per wave
  static constexpr int taps = KH * KW;
  static constexpr int NQ = (taps * CG + 3) / 4;  // 32-wide reduction chunks
  static constexpr int NP = 16 * NT;
  static constexpr int AR = kTH + KH - 1, AC = TW + KW - 1;
  static constexpr int a_vec = AR * AC * CG;      // 16-byte vectors of the input tile
  static constexpr int a_bytes = a_vec * 16;
  static constexpr int chunk_bytes = NP * RPB * 2;
  static constexpr int QG_max = (LDS_BUDGET - a_bytes) / chunk_bytes;
  static constexpr int NG = (NQ + QG_max - 1) / QG_max;  // weight groups per tile
  static constexpr int QG = (NQ + NG - 1) / NG;          // chunks per group (balanced)
  static constexpr int B_VPT = (QG * chunk_bytes / 16 + 255) / 256;
  static constexpr int lds_bytes = a_bytes + QG * chunk_bytes;
  static constexpr int out_bytes = kTH * TW * Cs * 2;
  static_assert(stage_count(CG) == 1, "the channel-staged layout has its own kernel");
  static_assert(QG_max >= 1, "conv tile does not fit in LDS");
  static_assert(out_bytes <= a_bytes, "output staging must fit in the input tile");
};

// One workgroup = 4 waves = a 4-row x TW-pixel output tile for all output
// channels.  The (4+KH-1) x (TW+KW-1) input halo tile is staged in LDS once;
// the packed weights stream through LDS in groups, the next group travelling
// in registers while the current one feeds the MFMAs.  Inside a group the A/B
// fragments of chunk q+1 are read from LDS while chunk q's MFMAs issue.
template <int KH, int KW, int CG, int NT, int TW, int LDS_BUDGET, int MINW>
__global__ __launch_bounds__(256, MINW) void k_conv_bf16(const bf16* __restrict__ x, int Hin, int Win,
                                                   const bf16* __restrict__ wpk, const float* __restrict__ bias,
                                                   int K, int act, bf16* __restrict__ y, int Ho, int Wo) {
  using Cfg = Bf16Cfg<KH, KW, CG, NT, TW, LDS_BUDGET>;
  constexpr int Cs = Cfg::Cs, MT = Cfg::MT;
  extern __shared__ __align__(16) unsigned char lds[];
  bf16* As = reinterpret_cast<bf16*>(lds);
  bf16* Bs = reinterpret_cast<bf16*>(lds + Cfg::a_bytes);
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int lrow = lane & 15, lgrp = lane >> 4;
  const int w0 = blockIdx.x * TW, h0 = blockIdx.y * kTH;

  // Weight group g: vectors [g*QG*chunk/16, ...) of the packed weights.
#define CCS_B_NVEC(g) (min(Cfg::QG, Cfg::NQ - (g) * Cfg::QG) * (Cfg::chunk_bytes / 16))
#define CCS_B_LOAD(g)                                                                                  \
  {                                                                                                    \
    const u32x4* src_ = reinterpret_cast<const u32x4*>(wpk) + (long)(g) * Cfg::QG * (Cfg::chunk_bytes / 16); \
    const int nv_ = CCS_B_NVEC(g);                                                                     \
    _Pragma("unroll") for (int i = 0; i < Cfg::B_VPT; ++i) wr[i] = src_[min(tid + 256 * i, nv_ - 1)]; \
  }
#define CCS_B_STORE(g)                                                                                 \
  {                                                                                                    \
    const int nv_ = CCS_B_NVEC(g);                                                                     \
    _Pragma("unroll") for (int i = 0; i < Cfg::B_VPT; ++i) {                                           \
      const int v_ = tid + 256 * i;                                                                    \
      if (v_ < nv_) reinterpret_cast<u32x4*>(Bs)[v_] = wr[i];                                          \
    }                                                                                                  \
  }

  u32x4 wr[Cfg::B_VPT];
  CCS_B_LOAD(0);
  {
    // All of this thread's input-tile vectors are loaded before any is stored,
    // so the tile costs one memory latency rather than one per vector.
    constexpr int vec_per_row = Cfg::AC * CG;
    constexpr int A_VPT = (Cfg::a_vec + 255) / 256;
    u32x4 ar[A_VPT];
#pragma unroll
    for (int i = 0; i < A_VPT; ++i) {
      const int v = min(tid + 256 * i, Cfg::a_vec - 1);
      const int r = v / vec_per_row, rem = v % vec_per_row;
      const int px = rem / CG, cg = rem % CG;
      const int hi = h0 + r, wi = w0 + px;
      const u32x4 val = *reinterpret_cast<const u32x4*>(
          x + (((long)min(hi, Hin - 1) * Win + min(wi, Win - 1)) * Cs + cg * 8));
      ar[i] = (hi < Hin && wi < Win) ? val : u32x4{0u, 0u, 0u, 0u};
    }
#pragma unroll
    for (int i = 0; i < A_VPT; ++i) {
      const int v = tid + 256 * i;
      if (v < Cfg::a_vec) reinterpret_cast<u32x4*>(As)[v] = ar[i];
    }
  }
  CCS_B_STORE(0);
  __syncthreads();

  // The MFMAs compute the transposed block (weights as the A operand): lane
  // (lrow, lgrp) accumulates pixel 16 m + lrow, channels 16 t + 4 lgrp .. +3,
  // so the epilogue stores four consecutive channels per lane at once.  The
  // accumulators start at the bias, so the epilogue has no add.
  f32x4 acc[MT][NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    f32x4 bn;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int n = 16 * t + 4 * lgrp + e;
      bn[e] = (n < K && bias) ? bias[n] : 0.f;
    }
#pragma unroll
    for (int m = 0; m < MT; ++m) acc[m][t] = bn;
  }

  auto load_frags = [&](int q, int ql, bf16x8 (&a)[MT], bf16x8 (&b)[NT]) {
    const int kg = 4 * q + lgrp;  // 8-element group of the reduction index
    const bool kvalid = kg < Cfg::taps * CG;
    const int tap = kvalid ? kg / CG : 0, c0 = kvalid ? (kg % CG) * 8 : 0;
    const int ti = tap / KW, tj = tap % KW;
#pragma unroll
    for (int m = 0; m < MT; ++m) {
      const bf16x8 v = *reinterpret_cast<const bf16x8*>(As + (((wave + ti) * Cfg::AC + 16 * m + lrow + tj) * Cs + c0));
      a[m] = kvalid ? v : bf16x8{};
    }
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const int n = 16 * t + lrow;
      b[t] = *reinterpret_cast<const bf16x8*>(Bs + ((ql * Cfg::NP + n) * RPB + 8 * bswz(n, lgrp)));
    }
  };

  for (int g = 0; g < Cfg::NG; ++g) {
    const bool more = g + 1 < Cfg::NG;
    if (more) CCS_B_LOAD(g + 1);
    const int q0 = g * Cfg::QG;
    const int nq = min(Cfg::QG, Cfg::NQ - q0);
    bf16x8 a[MT], b[NT];
    load_frags(q0, 0, a, b);
    for (int ql = 0; ql < nq; ++ql) {
      bf16x8 an[MT], bn[NT];
      if (ql + 1 < nq) load_frags(q0 + ql + 1, ql + 1, an, bn);
#if CCS_CONV_SCHED
      // Issue all of chunk q+1's LDS reads before chunk q's MFMAs, so their
      // latency hides behind the MFMAs instead of being scheduled just in time.
      __builtin_amdgcn_sched_barrier(0);
#endif
#pragma unroll
      for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int m = 0; m < MT; ++m) acc[m][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(b[t], a[m], acc[m][t], 0, 0, 0);
#if CCS_CONV_SCHED
      __builtin_amdgcn_sched_barrier(0);
#endif
#pragma unroll
      for (int m = 0; m < MT; ++m) a[m] = an[m];
#pragma unroll
      for (int t = 0; t < NT; ++t) b[t] = bn[t];
    }
    __syncthreads();  // every wave is done with this group (and, after the last one, with the tile)
    if (more) {
      CCS_B_STORE(g + 1);
      __syncthreads();
    }
  }

  // Epilogue: tanh -> bf16 staged in the input-tile region, then contiguous
  // stores.  The activation is chosen once, outside the element loops.  The
  // padding channels K <= n < Cs come out as exact zeros: their packed
  // weights and bias are zero, and tanh(0) = 0.
  bf16* Os = As;
  auto epilogue = [&](auto f) {
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const int n0 = 16 * t + 4 * lgrp;  // Cs is a multiple of 8: a group of 4 is all in or all out
#pragma unroll
      for (int m = 0; m < MT; ++m) {
        const int px = 16 * m + lrow;
        typedef __attribute__((ext_vector_type(4))) __bf16 bf16x4;
        const f32x4 r = {f(acc[m][t][0]), f(acc[m][t][1]), f(acc[m][t][2]), f(acc[m][t][3])};
        const bf16x4 v = __builtin_convertvector(r, bf16x4);  // two v_cvt_pk_bf16_f32
        if (n0 < Cs) *reinterpret_cast<bf16x4*>(Os + (wave * TW + px) * Cs + n0) = v;
      }
    }
  };
  if (act)
    epilogue([](float v) { return tanh_for_bf16(v); });  // the result is rounded to bf16
  else
    epilogue([](float v) { return v; });
  __syncthreads();
  constexpr int vec_per_row = TW * CG;
  for (int v = tid; v < kTH * vec_per_row; v += 256) {
    const int r = v / vec_per_row, rem = v % vec_per_row;
    const int px = rem / CG, cg = rem % CG;
    const int ho = h0 + r, wo = w0 + px;
    if (ho < Ho && wo < Wo)
      *reinterpret_cast<u32x4*>(y + (((long)ho * Wo + wo) * Cs + cg * 8)) =
          *reinterpret_cast<const u32x4*>(Os + ((r * TW + px) * Cs + cg * 8));
  }
}
#undef CCS_B_NVEC
#undef CCS_B_LOAD
#undef CCS_B_STORE

// ---------------------------------------------------------------------------
// BF16 MFMA implicit GEMM for the 100-channel layers (13 channel groups, two
// channel stages of the packed layout).
//
// A workgroup of 4 waves computes a 4-row x 64-pixel output tile for all 112
// (padded) output channels; wave w owns output row w: four 16-pixel M-tiles x
// seven 16-channel N-tiles of accumulators.  Per 32-deep reduction chunk a wave
// reads 4 input + 7 weight fragments of 1 KB from LDS for 28 MFMAs (the
// 32-pixel tile of k_conv_bf16 reads 2 + 7 for 14).  The (4+KH-1) x (64+KW-1)
// input halo tile is staged one channel stage at a time (groups 0-6, then 7-12,
// at a 112-byte pixel pitch, 44 KB), and the weights stream through a ring of
// four 8 KB chunk slots filled by direct global->LDS loads, so two 76 KB
// workgroups share a CU and no staging passes through registers.
//
// Ring protocol (chunk c in slot c % 4): the loads of chunk c are issued after
// the barrier that ends iteration c-4 (all reads of chunk c-4's slot are done by
// then); every wave waits for its loads of chunk c+2 (vmcnt: only chunk c+3's
// two loads may be younger) before the barrier that ends iteration c, so chunk
// c+2 is complete for every wave when iteration c+1 reads its fragments.
template <int KH, int KW>
struct StCfg {
  static constexpr int CG = 13, NT = 7, NP = 16 * NT, TW = 64, MT = TW / 16;
  static constexpr int taps = KH * KW;
  static constexpr int PG = stage_ng(CG, 0);  // 16-byte groups per LDS pixel (the larger stage)
  static constexpr int NQ0 = stage_nq(CG, taps, 0), NQ = packed_chunks(CG, taps);
  static constexpr int AR = kTH + KH - 1, AC = TW + KW - 1;
  static constexpr int a_vec = AR * AC * PG;
  static constexpr int A_RND = (a_vec + 255) / 256;  // direct loads per wave per stage
  static constexpr int a_bytes = A_RND * 256 * 16;   // the tile, padded to whole rounds
  static constexpr int chunk_vec = NP * RPB * 2 / 16;  // 448 vectors of 16 bytes
  static constexpr int slot_bytes = 8192;              // two loads of 64 vectors per wave
  static constexpr int NSLOT = 4;
  static constexpr int zero_off = a_bytes + NSLOT * slot_bytes;  // 16 zero bytes (padding groups read them)
  static constexpr int lds_bytes = zero_off + 16;
  static constexpr int Cs = CG * 8;
  static_assert(chunk_vec <= slot_bytes / 16, "a chunk fits its slot");
};

// Direct global->LDS load of 16 bytes per lane: lane i of the wave writes
// dst + 16 i (dst is wave-uniform).
__device__ __forceinline__ void load_lds16(const void* src, void* dst) {
  __builtin_amdgcn_global_load_lds(src, (__attribute__((address_space(3))) void*)dst, 16, 0, 0);
}

// s_waitcnt with vmcnt = N and the other counters left alone (gfx9 encoding:
// vmcnt[3:0] | expcnt[6:4] | lgkmcnt[11:8] | vmcnt[5:4] at [15:14]).
template <int N>
__device__ __forceinline__ void wait_vmcnt() {
  static_assert(N >= 0 && N < 64, "vmcnt range");
  __builtin_amdgcn_s_waitcnt((N & 15) | (7 << 4) | (15 << 8) | ((N >> 4) << 14));
}

// Barrier after which the LDS written before it (by direct loads whose vmcnt
// the caller has waited for, or by stores) is visible to every wave.
__device__ __forceinline__ void tile_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

template <int KH, int KW>
__global__ __launch_bounds__(256, 2) void k_conv_bf16_st(const bf16* __restrict__ x, int Hin, int Win,
                                                        const bf16* __restrict__ wpk, const float* __restrict__ bias,
                                                        int K, int act, bf16* __restrict__ y, int Ho, int Wo) {
  using Cfg = StCfg<KH, KW>;
  constexpr int Cs = Cfg::Cs, MT = Cfg::MT, NT = Cfg::NT, PG = Cfg::PG, AC = Cfg::AC;
  extern __shared__ __align__(16) unsigned char lds[];
  unsigned char* As = lds;
  unsigned char* Bs = lds + Cfg::a_bytes;
  const int tid = threadIdx.x;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
  const int lrow = lane & 15, lgrp = lane >> 4;
  const int w0 = blockIdx.x * Cfg::TW, h0 = blockIdx.y * kTH;
  const u32x4* wsrc = reinterpret_cast<const u32x4*>(wpk);
  constexpr int wvec_last = Cfg::NQ * Cfg::chunk_vec - 1;

  // Chunk c's weights into slot c % 4: this wave's two rows of 64 vectors
  // (vectors past the chunk's 448 read the next chunk, clamped to the buffer,
  // into the slot's padding).
  auto load_chunk = [&](int c) {
    if (c < Cfg::NQ) {
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        const int v = (2 * wave + i) * 64 + lane;
        load_lds16(wsrc + min(c * Cfg::chunk_vec + v, wvec_last),
                   Bs + (c % Cfg::NSLOT) * Cfg::slot_bytes + (2 * wave + i) * 1024);
      }
    }
  };
  // Stage st's input tile: vector v = (row, pixel, group) of the stage, pixel
  // pitch PG groups; addresses clamped to the map (the clamped values feed only
  // output pixels past Ho x Wo), vectors past the tile land in its padding.
  auto load_tile = [&](int st) {
    const int cg0 = stage_cg0(Cfg::CG, st);
#pragma unroll
    for (int i = 0; i < Cfg::A_RND; ++i) {
      const int v = min((i * 4 + wave) * 64 + lane, Cfg::a_vec - 1);
      const int r = v / (AC * PG), px = (v / PG) % AC, g = v % PG;
      const int hi = min(h0 + r, Hin - 1), wi = min(w0 + px, Win - 1);
      const bf16* src = x + (((long)hi * Win + wi) * Cs + min(cg0 + g, Cfg::CG - 1) * 8);
      load_lds16(src, As + (i * 4 + wave) * 1024);
    }
  };

#pragma unroll
  for (int c = 0; c < Cfg::NSLOT; ++c) load_chunk(c);
  load_tile(0);
  if (tid == 0) *reinterpret_cast<u32x4*>(lds + Cfg::zero_off) = u32x4{0u, 0u, 0u, 0u};

  // Accumulators start at the bias (transposed block: lane (lrow, lgrp) holds
  // pixel 16 m + lrow, channels 16 t + 4 lgrp .. +3).
  f32x4 acc[MT][NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    f32x4 bn;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int n = 16 * t + 4 * lgrp + e;
      bn[e] = (n < K && bias) ? bias[n] : 0.f;
    }
#pragma unroll
    for (int m = 0; m < MT; ++m) acc[m][t] = bn;
  }
  wait_vmcnt<0>();
  tile_barrier();

  // Fragments of chunk q (stage NGS groups, first chunk q0).  A padding group
  // reads the zero vector: the choice is made on the address, so nothing waits
  // for the loaded values before they feed the next chunk's MFMAs.
  auto load_frags = [&](auto ngc, int q, int q0, bf16x8 (&a)[MT], bf16x8 (&b)[NT]) {
    constexpr int NGS = decltype(ngc)::value;
    const int kg = 4 * (q - q0) + lgrp;
    const bool kvalid = kg < Cfg::taps * NGS;
    const int tap = kvalid ? kg / NGS : 0, g = kvalid ? kg % NGS : 0;
    const int ti = tap / KW, tj = tap % KW;
#pragma unroll
    for (int m = 0; m < MT; ++m) {
      const int off = (((wave + ti) * AC + 16 * m + lrow + tj) * PG + g) * 16;
      a[m] = *reinterpret_cast<const bf16x8*>(lds + (kvalid ? off : Cfg::zero_off));
    }
    const unsigned char* slot = Bs + (q % Cfg::NSLOT) * Cfg::slot_bytes;
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const int n = 16 * t + lrow;
      b[t] = *reinterpret_cast<const bf16x8*>(slot + ((n * RPB + 8 * bswz(n, lgrp)) * 2));
    }
  };

  // One chunk: chunk q+1's fragments into (an, bn) (unless q is the stage's
  // last), chunk q's MFMAs from (a, b), then the ring step.
  auto chunk_step = [&](auto ngc, int q, int q0, int q1, bf16x8 (&a)[MT], bf16x8 (&b)[NT], bf16x8 (&an)[MT],
                        bf16x8 (&bn)[NT]) {
    if (q + 1 < q1) load_frags(ngc, q + 1, q0, an, bn);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
      for (int m = 0; m < MT; ++m) acc[m][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(b[t], a[m], acc[m][t], 0, 0, 0);
    __builtin_amdgcn_sched_barrier(0);
    // This wave's loads of chunk q+2 are done: only chunk q+3's two loads may
    // be younger, and there are none when q+3 is past the last chunk.
    if (q + 3 < Cfg::NQ)
      wait_vmcnt<2>();
    else
      wait_vmcnt<0>();
    tile_barrier();
    load_chunk(q + Cfg::NSLOT);
  };
  // One channel stage: chunks q0 .. q1-1, two per iteration with the fragment
  // registers swapping roles (no copies between chunks).
  auto run_stage = [&](auto ngc, int q0, int q1) {
    bf16x8 a0[MT], b0[NT], a1[MT], b1[NT];
    load_frags(ngc, q0, q0, a0, b0);
    int q = q0;
    for (; q + 1 < q1; q += 2) {
      chunk_step(ngc, q, q0, q1, a0, b0, a1, b1);
      chunk_step(ngc, q + 1, q0, q1, a1, b1, a0, b0);
    }
    if (q < q1) chunk_step(ngc, q, q0, q1, a0, b0, a1, b1);
  };
  run_stage(std::integral_constant<int, stage_ng(13, 0)>{}, 0, Cfg::NQ0);
  load_tile(1);  // every wave has passed the barrier after its last stage-0 read
  wait_vmcnt<0>();
  tile_barrier();
  run_stage(std::integral_constant<int, stage_ng(13, 1)>{}, Cfg::NQ0, Cfg::NQ);
  wait_vmcnt<0>();

  // Epilogue: + tanh, bf16, stored straight from the accumulators: a lane
  // holds four consecutive channels of one pixel (8 bytes); the seven N-tiles
  // of a pixel fill its 208-byte row, which the L2 merges before writing back.
  // The padding channels K <= n < Cs come out as exact zeros (zero weights and
  // bias, tanh(0) = 0).
  const int ho = h0 + wave;
  auto epilogue = [&](auto f) {
    if (ho >= Ho) return;
    bf16* yrow = y + (long)ho * Wo * Cs;
#pragma unroll
    for (int m = 0; m < MT; ++m) {
      const int wo = w0 + 16 * m + lrow;
      if (wo >= Wo) continue;
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        const int n0 = 16 * t + 4 * lgrp;
        typedef __attribute__((ext_vector_type(4))) __bf16 bf16x4;
        const f32x4 r = {f(acc[m][t][0]), f(acc[m][t][1]), f(acc[m][t][2]), f(acc[m][t][3])};
        const bf16x4 v = __builtin_convertvector(r, bf16x4);
        if (n0 < Cs) *reinterpret_cast<bf16x4*>(yrow + (long)wo * Cs + n0) = v;
      }
    }
  };
  if (act)
    epilogue([](float v) { return tanh_for_bf16(v); });
  else
    epilogue([](float v) { return v; });
}

// ---------------------------------------------------------------------------
// FP32 MFMA implicit GEMM (v_mfma_f32_16x16x4_f32)

constexpr int kTHf = 2;
constexpr int kTWf = 64;

template <int KH, int KW, int CG, int NT>
struct F32Cfg {
  static constexpr int Cs = CG * 8;
  static constexpr int CP = Cs + 1;  // LDS pixel pitch in floats (odd: conflict-free)
  static constexpr int taps = KH * KW;
  static constexpr int NQ = (taps * Cs + 3) / 4;
  static constexpr int NP = 16 * NT;
  static constexpr int AR = kTHf + KH - 1, AC = kTWf + KW - 1;
  static constexpr int a_bytes = (AR * AC * CP * 4 + 15) / 16 * 16;
  static constexpr int chunk_bytes = 4 * NP * 4;
  static constexpr int QG_max = (kLdsBytes - a_bytes) / chunk_bytes;
  static constexpr int QG = QG_max < NQ ? QG_max : NQ;
  static constexpr int out_bytes = kTHf * kTWf * Cs * 4;
  static_assert(QG >= 1, "conv tile does not fit in LDS");
  static_assert(out_bytes <= a_bytes, "output staging must fit in the input tile");
};

template <int KH, int KW, int CG, int NT>
__global__ __launch_bounds__(256) void k_conv_f32(const float* __restrict__ x, int Hin, int Win,
                                                  const float* __restrict__ wpk, const float* __restrict__ bias,
                                                  int K, int act, float* __restrict__ y, int Ho, int Wo) {
  using Cfg = F32Cfg<KH, KW, CG, NT>;
  constexpr int Cs = Cfg::Cs, CP = Cfg::CP;
  extern __shared__ __align__(16) unsigned char lds[];
  float* As = reinterpret_cast<float*>(lds);
  float* Bs = reinterpret_cast<float*>(lds + Cfg::a_bytes);

  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int wrow = wave >> 1, wcol = (wave & 1) * 32;
  const int w0 = blockIdx.x * kTWf, h0 = blockIdx.y * kTHf;
  {
    // 16-byte loads (Cs is a multiple of 8), scalar stores into the odd LDS pitch.
    constexpr int vpp = Cs / 4;  // float4 per pixel
    constexpr int per_row = Cfg::AC * vpp;
    batched_copy16<9>(
        Cfg::AR * per_row, tid, 256,
        [&](int v) {
          const int r = v / per_row, rem = v % per_row;
          const int px = rem / vpp, c4 = rem % vpp;
          const int hi = h0 + r, wi = w0 + px;
          const u32x4 val = *reinterpret_cast<const u32x4*>(
              x + ((long)min(hi, Hin - 1) * Win + min(wi, Win - 1)) * Cs + 4 * c4);
          return (hi < Hin && wi < Win) ? val : u32x4{0u, 0u, 0u, 0u};
        },
        [&](int v, const u32x4& val) {
          const int r = v / per_row, rem = v % per_row;
          const int px = rem / vpp, c4 = rem % vpp;
          float* d = As + (r * Cfg::AC + px) * CP + 4 * c4;
#pragma unroll
          for (int e = 0; e < 4; ++e) d[e] = __uint_as_float(val[e]);
        });
  }
  f32x4 acc[2][NT];
#pragma unroll
  for (int m = 0; m < 2; ++m)
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[m][t] = f32x4{0.f, 0.f, 0.f, 0.f};

  const int lrow = lane & 15, lk = lane >> 4;
  for (int q0 = 0; q0 < Cfg::NQ; q0 += Cfg::QG) {
    const int nq = (Cfg::NQ - q0) < Cfg::QG ? (Cfg::NQ - q0) : Cfg::QG;
    __syncthreads();
    {
      const u32x4* src = reinterpret_cast<const u32x4*>(wpk + (long)q0 * 4 * Cfg::NP);
      u32x4* dst = reinterpret_cast<u32x4*>(Bs);
      batched_copy16<8>(nq * Cfg::chunk_bytes / 16, tid, 256, [&](int v) { return src[v]; },
                        [&](int v, const u32x4& val) { dst[v] = val; });
    }
    __syncthreads();
    for (int ql = 0; ql < nq; ++ql) {
      const int k = 4 * (q0 + ql) + lk;
      const int tap = k / Cs, c = k % Cs;
      const bool kvalid = tap < Cfg::taps;
      const int ti = kvalid ? tap / KW : 0, tj = kvalid ? tap % KW : 0;
      float a[2];
#pragma unroll
      for (int m = 0; m < 2; ++m)
        a[m] = kvalid ? As[((wrow + ti) * Cfg::AC + wcol + 16 * m + lrow + tj) * CP + c] : 0.f;
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        const float b = Bs[(ql * 4 + lk) * Cfg::NP + 16 * t + lrow];
#pragma unroll
        for (int m = 0; m < 2; ++m) acc[m][t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[m], b, acc[m][t], 0, 0, 0);
      }
    }
  }
  __syncthreads();
  float* Os = As;
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const int n = 16 * t + lrow;
    const float bn = (n < K && bias) ? bias[n] : 0.f;
#pragma unroll
    for (int m = 0; m < 2; ++m)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int px = wcol + 16 * m + 4 * (lane >> 4) + e;
        float v = acc[m][t][e] + bn;
        if (act) v = tanh_fast(v);
        if (n < Cs) Os[(wrow * kTWf + px) * Cs + n] = n < K ? v : 0.f;
      }
  }
  __syncthreads();
  {
    constexpr int vec_per_row = kTWf * Cs / 4;
    for (int v = tid; v < kTHf * vec_per_row; v += 256) {
      const int r = v / vec_per_row, rem = v % vec_per_row;
      const int px = (rem * 4) / Cs;
      const int ho = h0 + r, wo = w0 + px;
      if (ho < Ho && wo < Wo)
        *reinterpret_cast<float4*>(y + ((long)ho * Wo + w0) * Cs + (long)rem * 4) =
            *reinterpret_cast<const float4*>(Os + (r * kTWf) * Cs + rem * 4);
    }
  }
}

// ---------------------------------------------------------------------------
// Generic VALU convolution: one thread per (output pixel, 8 output channels).
// Input either channel-last (stride Cs_in, bf16 or f32) or C x H x W float32.

template <typename TI, typename TO>
__global__ void k_conv_generic(const TI* __restrict__ x, int x_nchw, int Hin, int Win, int C, int Cs_in,
                               const float* __restrict__ w, const float* __restrict__ bias, int K, int KH, int KW,
                               int act, TO* __restrict__ y, int Ho, int Wo, int Cs_out) {
  const int groups = Cs_out / 8;
  const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (long)Ho * Wo * groups) return;
  const int g = (int)(gid % groups);
  const long p = gid / groups;
  const int ho = (int)(p / Wo), wo = (int)(p % Wo);
  float acc[8];
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const int n = g * 8 + e;
    acc[e] = (n < K && bias) ? bias[n] : 0.f;
  }
  for (int i = 0; i < KH; ++i)
    for (int j = 0; j < KW; ++j)
      for (int c = 0; c < C; ++c) {
        const long xi = x_nchw ? ((long)c * Hin + ho + i) * Win + wo + j
                               : ((long)(ho + i) * Win + wo + j) * Cs_in + c;
        const float xv = (float)x[xi];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          const int n = g * 8 + e;
          if (n < K) acc[e] += xv * w[(((long)n * C + c) * KH + (KH - 1 - i)) * KW + (KW - 1 - j)];
        }
      }
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const int n = g * 8 + e;
    float v = acc[e];
    if (act) v = sizeof(TO) == 2 ? tanh_for_bf16(v) : tanh_fast(v);
    y[p * Cs_out + n] = (TO)(n < K ? v : 0.f);
  }
}

// ---------------------------------------------------------------------------
// Image layer: C <= 5 input channels (C x H x W float32, the normalised padded
// image), 3x3, K <= 128 outputs.  One thread per output pixel computes every
// output channel; the flipped weights sit in LDS and the block's 256 x Cs output
// tile is staged in LDS for contiguous stores.

template <int C, typename TO>
__global__ __launch_bounds__(256) void k_conv_image(const float* __restrict__ x, int Hin, int Win,
                                                    const float* __restrict__ w, const float* __restrict__ bias,
                                                    int K, int act, TO* __restrict__ y, int Ho, int Wo) {
  constexpr int KMAX = 128;
  // FP32 outputs (32 bytes per 8 channels) go straight to memory; bf16 outputs
  // (16 bytes) are staged through LDS and copied out as one contiguous block,
  // which measured faster for them (0.051 vs 0.053 ms at 1256 x 389; FP32:
  // 0.080 ms direct, 0.109 ms staged).
  constexpr bool kDirect = sizeof(TO) == 4;
  __shared__ float ws[C * 9][KMAX];
  __shared__ float bs[KMAX];
  extern __shared__ __align__(16) unsigned char lds_out[];
  TO* os = reinterpret_cast<TO*>(lds_out);
  const int Cs = (K + 7) / 8 * 8;
  for (int v = threadIdx.x; v < C * 9 * KMAX; v += 256) {
    const int k = v % KMAX, t = v / KMAX;  // t = c * 9 + i * 3 + j
    const int c = t / 9, i = (t % 9) / 3, j = t % 3;
    ws[t][k] = k < K ? w[(((long)k * C + c) * 3 + (2 - i)) * 3 + (2 - j)] : 0.f;
  }
  for (int k = threadIdx.x; k < KMAX; k += 256) bs[k] = (k < K && bias) ? bias[k] : 0.f;
  __syncthreads();
  const long p0 = (long)blockIdx.x * 256;
  const long p = p0 + threadIdx.x;
  const long P = (long)Ho * Wo;
  float in[C * 9];
  if (p < P) {
    const int ho = (int)(p / Wo), wo = (int)(p % Wo);
#pragma unroll
    for (int c = 0; c < C; ++c)
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) in[c * 9 + i * 3 + j] = x[((long)c * Hin + ho + i) * Win + wo + j];
  } else {
#pragma unroll
    for (int t = 0; t < C * 9; ++t) in[t] = 0.f;
  }
  for (int k0 = 0; k0 < Cs; k0 += 8) {
    float acc[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) acc[e] = bs[k0 + e];
#pragma unroll
    for (int t = 0; t < C * 9; ++t) {
      const float4 w0 = *reinterpret_cast<const float4*>(&ws[t][k0]);
      const float4 w1 = *reinterpret_cast<const float4*>(&ws[t][k0 + 4]);
      acc[0] += in[t] * w0.x; acc[1] += in[t] * w0.y; acc[2] += in[t] * w0.z; acc[3] += in[t] * w0.w;
      acc[4] += in[t] * w1.x; acc[5] += in[t] * w1.y; acc[6] += in[t] * w1.z; acc[7] += in[t] * w1.w;
    }
    // Eight channels, rounded to TO, as one vector (16 bytes in bf16).
    typedef __attribute__((ext_vector_type(8))) TO vec8;
    vec8 o;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      float v = acc[e];
      if (act) v = sizeof(TO) == 2 ? tanh_for_bf16(v) : tanh_fast(v);
      o[e] = (TO)(k0 + e < K ? v : 0.f);
    }
    if constexpr (kDirect) {
      if (p < P) *reinterpret_cast<vec8*>(y + p * Cs + k0) = o;
    } else {
      *reinterpret_cast<vec8*>(os + threadIdx.x * Cs + k0) = o;
    }
  }
  if constexpr (!kDirect) {
    __syncthreads();
    // The block's pixels are consecutive in memory: copy 16-byte vectors.
    const long npx = (P - p0) < 256 ? (P - p0) : 256;
    const int nvec = (int)(npx * Cs * sizeof(TO) / 16);
    const uint4* src = reinterpret_cast<const uint4*>(os);
    uint4* dst = reinterpret_cast<uint4*>(y + p0 * Cs);
    for (int v = threadIdx.x; v < nvec; v += 256) dst[v] = src[v];
  }
}

// ---------------------------------------------------------------------------
// Dispatch

struct ConvShape {
  int H, W, C, K, KH, KW;
};

// Tile width, LDS budget per workgroup and minimum waves per SIMD of
// k_conv_bf16, per filter size (chosen when it also ran the 100-channel layers,
// measured on 1242 x 375 x 100-channel layers: 3x3 -- 32-pixel tiles, two 80 KB
// workgroups per CU, 0.171 ms per layer, 16-pixel tiles at three waves per
// SIMD: 0.192 ms; 2x2 -- 16-pixel tiles, three 52000-byte workgroups per CU,
// 0.109 ms, 32-pixel tiles: 0.123 ms).
template <int KH>
struct Bf16Tile {
  static constexpr int TW = 32, LDS = 80 * 1024, MINW = 1;
};
template <>
struct Bf16Tile<2> {
  static constexpr int TW = 16, LDS = 52000, MINW = 3;
};

template <int KH, int KW>
static bool try_bf16_st(const ConvShape& s, const void* x, const void* wpk, const float* b, int act, void* y,
                        hipStream_t st) {
  if (s.KH != KH || s.KW != KW || feature_stride(s.C) != 104 || (s.K + 15) / 16 != 7) return false;
  using Cfg = StCfg<KH, KW>;
  const auto kern = k_conv_bf16_st<KH, KW>;
  const int Ho = s.H - KH + 1, Wo = s.W - KW + 1;
  static bool configured = false;
  if (!configured) {
    (void)hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, Cfg::lds_bytes);
    configured = true;
  }
  kern<<<dim3(ceil_div(Wo, Cfg::TW), ceil_div(Ho, kTH)), 256, Cfg::lds_bytes, st>>>(
      (const bf16*)x, s.H, s.W, (const bf16*)wpk, b, s.K, act, (bf16*)y, Ho, Wo);
  return true;
}

template <int KH, int KW, int CG, int NT>
static bool try_bf16(const ConvShape& s, const void* x, const void* wpk, const float* b, int act, void* y,
                     hipStream_t st) {
  if (s.KH != KH || s.KW != KW || feature_stride(s.C) != CG * 8 || (s.K + 15) / 16 != NT) return false;
  using T = Bf16Tile<KH>;
  using Cfg = Bf16Cfg<KH, KW, CG, NT, T::TW, T::LDS>;
  const auto kern = k_conv_bf16<KH, KW, CG, NT, T::TW, T::LDS, T::MINW>;
  const int Ho = s.H - KH + 1, Wo = s.W - KW + 1;
  static bool configured = false;
  if (!configured) {
    (void)hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, kLdsBytes);
    configured = true;
  }
  kern<<<dim3(ceil_div(Wo, T::TW), ceil_div(Ho, kTH)), 256, Cfg::lds_bytes, st>>>(
      (const bf16*)x, s.H, s.W, (const bf16*)wpk, b, s.K, act, (bf16*)y, Ho, Wo);
  return true;
}

template <int KH, int KW, int CG, int NT>
static bool try_f32(const ConvShape& s, const void* x, const void* wpk, const float* b, int act, void* y,
                    hipStream_t st) {
  if (s.KH != KH || s.KW != KW || feature_stride(s.C) != CG * 8 || (s.K + 15) / 16 != NT) return false;
  using Cfg = F32Cfg<KH, KW, CG, NT>;
  const int Ho = s.H - KH + 1, Wo = s.W - KW + 1;
  const size_t lds = Cfg::a_bytes + (size_t)Cfg::QG * Cfg::chunk_bytes;
  static bool configured = false;
  if (!configured) {
    (void)hipFuncSetAttribute((const void*)k_conv_f32<KH, KW, CG, NT>, hipFuncAttributeMaxDynamicSharedMemorySize,
                        kLdsBytes);
    configured = true;
  }
  k_conv_f32<KH, KW, CG, NT><<<dim3(ceil_div(Wo, kTWf), ceil_div(Ho, kTHf)), 256, lds, st>>>(
      (const float*)x, s.H, s.W, (const float*)wpk, b, s.K, act, (float*)y, Ho, Wo);
  return true;
}

// MFMA kernels exist for the 100-channel layers of both tower architectures
// (3x3 and 2x2), the 64-channel 3x3 layer of the pairwise network
// (stereo/stereonet.cpp:158-164) and 8-channel layers (small test networks).
// They write the output with the input's channel stride, so both strides must
// agree.
static bool mfma_supported(const ConvShape& s) {
  const int cs = feature_stride(s.C), nt = (s.K + 15) / 16;
  const bool k3 = s.KH == 3 && s.KW == 3, k2 = s.KH == 2 && s.KW == 2;
  if (!(k3 || k2) || feature_stride(s.K) != cs) return false;
  return (cs == 104 && nt == 7) || (cs == 8 && nt == 1) || (k3 && cs == 64 && nt == 4);
}

static bool launch_mfma(const ConvShape& s, int precision, const void* x, const void* wpk, const float* b,
                        int act, void* y, hipStream_t st) {
  if (precision == CCS_BF16) {
    return try_bf16_st<3, 3>(s, x, wpk, b, act, y, st) || try_bf16_st<2, 2>(s, x, wpk, b, act, y, st) ||
           try_bf16<3, 3, 8, 4>(s, x, wpk, b, act, y, st) || try_bf16<3, 3, 1, 1>(s, x, wpk, b, act, y, st) ||
           try_bf16<2, 2, 1, 1>(s, x, wpk, b, act, y, st);
  }
  return try_f32<3, 3, 13, 7>(s, x, wpk, b, act, y, st) || try_f32<2, 2, 13, 7>(s, x, wpk, b, act, y, st) ||
         try_f32<3, 3, 8, 4>(s, x, wpk, b, act, y, st) || try_f32<3, 3, 1, 1>(s, x, wpk, b, act, y, st) ||
         try_f32<2, 2, 1, 1>(s, x, wpk, b, act, y, st);
}

static size_t packed_bytes(int K, int C, int KH, int KW, int precision) {
  const int Cs = feature_stride(C);
  const int NP = 16 * ((K + 15) / 16);
  if (precision == CCS_BF16) {
    const int NQ = packed_chunks(Cs / 8, KH * KW);
    return (size_t)NQ * NP * RPB * 2;
  }
  const int NQ = (KH * KW * Cs + 3) / 4;
  return (size_t)NQ * 4 * NP * 4;
}

}  // namespace ccs

using namespace ccs;

extern "C" size_t ccs_conv_packed_bytes(int K, int C, int KH, int KW, int precision) {
  return packed_bytes(K, C, KH, KW, precision);
}

extern "C" int ccs_conv_pack_weights(ccs_handle_t h, const float* w, int K, int C, int KH, int KW, int precision,
                                     void* packed) {
  CCS_CHECK(h && w && packed, CCS_ERR_ARG, "ccs_conv_pack_weights: NULL argument");
  CCS_CHECK(K >= 1 && C >= 1 && KH >= 1 && KW >= 1, CCS_ERR_ARG, "ccs_conv_pack_weights: bad shape");
  CCS_CHECK(precision == CCS_FP32 || precision == CCS_BF16, CCS_ERR_ARG, "ccs_conv_pack_weights: bad precision");
  const int Cs = feature_stride(C);
  const int NP = 16 * ((K + 15) / 16);
  if (precision == CCS_BF16) {
    const int NQ = packed_chunks(Cs / 8, KH * KW);
    const long total = (long)NQ * NP * RPB;
    k_pack_bf16<<<ceil_div(total, 256), 256, 0, h->stream>>>(w, K, C, Cs, KH, KW, NQ, NP, (bf16*)packed);
  } else {
    const int NQ = (KH * KW * Cs + 3) / 4;
    const long total = (long)NQ * 4 * NP;
    k_pack_f32<<<ceil_div(total, 256), 256, 0, h->stream>>>(w, K, C, Cs, KH, KW, NQ, NP, (float*)packed);
  }
  CCS_LAUNCH_CHECK();
  return CCS_OK;
}

template <int C>
static void launch_image(const float* x, int H, int W, const float* w, const float* b, int K, int act,
                         int precision, void* y, hipStream_t st) {
  const int Ho = H - 2, Wo = W - 2;
  const long P = (long)Ho * Wo;
  const int Cs = feature_stride(K);
  if (precision == CCS_BF16)  // bf16 outputs are staged in LDS (256 pixels x Cs)
    k_conv_image<C, bf16><<<ceil_div(P, 256), 256, 256 * Cs * 2, st>>>(x, H, W, w, b, K, act, (bf16*)y, Ho, Wo);
  else
    k_conv_image<C, float><<<ceil_div(P, 256), 256, 0, st>>>(x, H, W, w, b, K, act, (float*)y, Ho, Wo);
}

static int conv_generic(ccs_handle_t h, const void* x, int x_layout, int H, int W, int C, const float* w,
                        const float* b, int K, int KH, int KW, int act, int precision, void* y) {
  const int Ho = H - KH + 1, Wo = W - KW + 1;
  const int Cs_out = feature_stride(K);
  if (x_layout == CCS_LAYOUT_NCHW && KH == 3 && KW == 3 && K <= 128 && C <= 5) {
    switch (C) {
      case 1: launch_image<1>((const float*)x, H, W, w, b, K, act, precision, y, h->stream); break;
      case 2: launch_image<2>((const float*)x, H, W, w, b, K, act, precision, y, h->stream); break;
      case 3: launch_image<3>((const float*)x, H, W, w, b, K, act, precision, y, h->stream); break;
      case 4: launch_image<4>((const float*)x, H, W, w, b, K, act, precision, y, h->stream); break;
      default: launch_image<5>((const float*)x, H, W, w, b, K, act, precision, y, h->stream); break;
    }
    CCS_LAUNCH_CHECK();
    return CCS_OK;
  }
  const long threads = (long)Ho * Wo * (Cs_out / 8);
  const unsigned grid = ceil_div(threads, 256);
  const int nchw = x_layout == CCS_LAYOUT_NCHW;
  const int Cs_in = nchw ? C : feature_stride(C);
  if (precision == CCS_BF16) {
    if (nchw)
      k_conv_generic<float, bf16><<<grid, 256, 0, h->stream>>>((const float*)x, 1, H, W, C, Cs_in, w, b, K, KH, KW,
                                                              act, (bf16*)y, Ho, Wo, Cs_out);
    else
      k_conv_generic<bf16, bf16><<<grid, 256, 0, h->stream>>>((const bf16*)x, 0, H, W, C, Cs_in, w, b, K, KH, KW,
                                                             act, (bf16*)y, Ho, Wo, Cs_out);
  } else {
    k_conv_generic<float, float><<<grid, 256, 0, h->stream>>>((const float*)x, nchw, H, W, C, Cs_in, w, b, K, KH,
                                                             KW, act, (float*)y, Ho, Wo, Cs_out);
  }
  CCS_LAUNCH_CHECK();
  return CCS_OK;
}

static int check_conv(const void* x, int x_layout, int H, int W, int C, int K, int KH, int KW, int precision,
                      const void* y) {
  CCS_CHECK(x && y, CCS_ERR_ARG, "ccs_conv2d: NULL tensor");
  CCS_CHECK(precision == CCS_FP32 || precision == CCS_BF16, CCS_ERR_ARG, "ccs_conv2d: bad precision %d", precision);
  CCS_CHECK(x_layout == CCS_LAYOUT_NHWC || x_layout == CCS_LAYOUT_NCHW, CCS_ERR_ARG, "ccs_conv2d: bad layout");
  CCS_CHECK(C >= 1 && K >= 1 && KH >= 1 && KW >= 1 && H >= KH && W >= KW, CCS_ERR_SHAPE,
            "ccs_conv2d: input %dx%d smaller than the %dx%d kernel (or empty)", H, W, KH, KW);
  CCS_CHECK(C <= 4096 && K <= 4096, CCS_ERR_UNSUPPORTED, "ccs_conv2d: too many channels");
  return CCS_OK;
}

extern "C" int ccs_conv_forward_packed(ccs_handle_t h, const void* x, int x_layout, int H, int W, int C,
                                       const void* packed, const float* b, int K, int KH, int KW, int act,
                                       int precision, void* y) {
  CCS_CHECK(h && packed, CCS_ERR_ARG, "ccs_conv_forward_packed: NULL argument");
  int rc = check_conv(x, x_layout, H, W, C, K, KH, KW, precision, y);
  if (rc) return rc;
  const ConvShape s{H, W, C, K, KH, KW};
  CCS_CHECK(x_layout == CCS_LAYOUT_NHWC && mfma_supported(s), CCS_ERR_UNSUPPORTED,
            "ccs_conv_forward_packed: no MFMA kernel for C=%d K=%d %dx%d (use ccs_conv2d_forward)", C, K, KH, KW);
  launch_mfma(s, precision, x, packed, b, act, y, h->stream);
  CCS_LAUNCH_CHECK();
  return CCS_OK;
}

// ---- allocation-free variant ----------------------------------------------
// The packed weights of ccs_conv2d_forward_noalloc live in one fixed array per
// GPU (held until the process ends) instead of the handle's scratch.  The
// array's name is a host-side stand-in, not a device address: host code uses
// only the address hipGetSymbolAddress returns on each device.
constexpr size_t kNoAllocScratchBytes = 512u << 10;  // the largest MFMA pack is 419,328 bytes
constexpr int kMaxDevices = 64;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices] = {};

// Resolves the current device's copy of `scratch` once (ccs_create calls this
// for the handle's device; the entry point repeats it on a device first used
// without a handle created there).
int ccs::noalloc_scratch_setup() {
  int dev = 0;
  CCS_HIP(hipGetDevice(&dev));
  CCS_CHECK(dev >= 0 && dev < kMaxDevices, CCS_ERR_UNSUPPORTED, "device %d beyond %d", dev, kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev]) return CCS_OK;
  void* p = nullptr;
  CCS_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  CCS_CHECK(p != nullptr, CCS_ERR_HIP, "hipGetSymbolAddress returned no address for the scratch array");
  g_scratch_ptr[dev] = p;
  return CCS_OK;
}

static int noalloc_scratch(void** out) {
  int dev = 0;
  CCS_HIP(hipGetDevice(&dev));
  CCS_CHECK(dev >= 0 && dev < kMaxDevices, CCS_ERR_UNSUPPORTED, "device %d beyond %d", dev, kMaxDevices);
  {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    *out = g_scratch_ptr[dev];
  }
  if (*out) return CCS_OK;
  int rc = noalloc_scratch_setup();
  if (rc) return rc;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  *out = g_scratch_ptr[dev];
  return CCS_OK;
}

extern "C" int ccs_conv2d_forward_noalloc(ccs_handle_t h, const void* x, int x_layout, int H, int W, int C,
                                          const float* w, const float* b, int K, int KH, int KW, int act,
                                          int precision, void* y) {
  CCS_CHECK(h && w, CCS_ERR_ARG, "ccs_conv2d_forward_noalloc: NULL argument");
  int rc = check_conv(x, x_layout, H, W, C, K, KH, KW, precision, y);
  if (rc) return rc;
  const ConvShape s{H, W, C, K, KH, KW};
  if (x_layout == CCS_LAYOUT_NHWC && mfma_supported(s)) {
    const size_t need = packed_bytes(K, C, KH, KW, precision);
    CCS_CHECK(need <= kNoAllocScratchBytes, CCS_ERR_UNSUPPORTED,
              "ccs_conv2d_forward_noalloc: %d -> %d %dx%d needs %zu bytes of packed weights, the fixed scratch "
              "holds %zu",
              C, K, KH, KW, need, kNoAllocScratchBytes);
    void* p = nullptr;
    if ((rc = noalloc_scratch(&p))) return rc;
    if ((rc = ccs_conv_pack_weights(h, w, K, C, KH, KW, precision, p))) return rc;
    launch_mfma(s, precision, x, p, b, act, y, h->stream);
    CCS_LAUNCH_CHECK();
    return CCS_OK;
  }
  CCS_CHECK(x_layout == CCS_LAYOUT_NHWC || C <= 8, CCS_ERR_UNSUPPORTED,
            "ccs_conv2d_forward_noalloc: NCHW input is only accepted for image layers (C <= 8)");
  return conv_generic(h, x, x_layout, H, W, C, w, b, K, KH, KW, act, precision, y);  // reads w directly
}

extern "C" int ccs_conv2d_forward(ccs_handle_t h, const void* x, int x_layout, int H, int W, int C, const float* w,
                                  const float* b, int K, int KH, int KW, int act, int precision, void* y) {
  CCS_CHECK(h && w, CCS_ERR_ARG, "ccs_conv2d_forward: NULL argument");
  int rc = check_conv(x, x_layout, H, W, C, K, KH, KW, precision, y);
  if (rc) return rc;
  const ConvShape s{H, W, C, K, KH, KW};
  if (x_layout == CCS_LAYOUT_NHWC && mfma_supported(s)) {
    if ((rc = h->conv_w.reserve(packed_bytes(K, C, KH, KW, precision), h->stream))) return rc;
    if ((rc = ccs_conv_pack_weights(h, w, K, C, KH, KW, precision, h->conv_w.ptr))) return rc;
    launch_mfma(s, precision, x, h->conv_w.ptr, b, act, y, h->stream);
    CCS_LAUNCH_CHECK();
    return CCS_OK;
  }
  CCS_CHECK(x_layout == CCS_LAYOUT_NHWC || C <= 8, CCS_ERR_UNSUPPORTED,
            "ccs_conv2d_forward: NCHW input is only accepted for image layers (C <= 8)");
  return conv_generic(h, x, x_layout, H, W, C, w, b, K, KH, KW, act, precision, y);
}
