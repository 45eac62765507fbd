// SlackProp CRF inference on MI355X: dual minorize-maximize over row and column
// chains of the 4-connected grid.
//
// The reference calls slack_prop_2D_alg from an un-vendored library
// (stereo/stereo.cpp:140-154, stereo/operators/slackprop.cpp:96-105); this file
// implements a restatement of that published method:
//
//   A = C                                   (row unaries; columns hold C - A)
//   repeat iters times:
//     rows:    g = minorant(A_row, wx_row);       A_row -= g
//     columns: g = minorant((C-A)_col, wy_col);   A_col += g
//   decode rows: label = argmin of the row min-marginals
//
// with the pairwise term w * rho(x_p - x_q), rho = 0 / L1 (0 < |d| < delta) /
// L2 (|d| >= delta).  minorant() is a tight, maximal modular minorant of one
// chain: a backward min-sum pass stores normalised messages Bh, then a forward
// pass keeps g_i = 1/4 (R_i + B_i) + excess_i at every node and passes the rest
// on.  The dual bound Phi_rows(A) + Phi_cols(C - A) never decreases; its float64
// bookkeeping (chain minima S, kept constants K) runs beside the float32 min-sum.
//
// The sweeps keep one state volume that alternates between the two halves'
// unaries: a row sweep reads A and stores B = C - (A - g), a column sweep reads
// B and stores A = (C - B) + g, so each reads one volume in each of its two
// passes and the cost volume only in its forward pass (4 volume passes per
// sweep instead of 3 for rows and 5 for columns).
//
// GPU mapping: one wavefront per chain, label s = lane * LPL + j held in VGPRs,
// DPP wave_shr/wave_shl for the label neighbours and a DPP tree for the minimum
// over labels.  The volume streams through HBM once per pass (node i is D
// contiguous floats), backward messages go to a scratch volume (column sweeps
// of large volumes store every seg_len()-th one and recompute the rest), and the
// per-node minima go to LDS; their float64 sums are formed by wave scans after
// each pass.  The row decode runs its two message passes on two wavefronts at
// once, and two more waves take its decisions (k_decode4).
//
// This file holds the chain sweeps only.  It is compiled with -fno-honor-nans:
// the min/max trees then run as fused DPP instructions without NaN-quieting
// moves.  Costs must be finite or +inf; a NaN cost gives unspecified labels
// (never an out-of-range memory access: labels are only stored, and every
// address is computed from chain/node/lane indices).  The row minorant can
// flag NaN unaries per chain (ChainArgs::nan_flags) with a class test of every
// value it reads, which the host turns into a status code.
#include <cstdlib>

#include "crf_chain.hpp"

namespace ccs {

constexpr float kInfF = __builtin_huge_valf();
// Relative tolerance under which two decoded min-marginals count as tied.
#ifndef CCS_CRF_TIE_EPS
#define CCS_CRF_TIE_EPS 1e-5f
#endif
constexpr float kTieEps = CCS_CRF_TIE_EPS;

// Scheduling fence between a ring slot's read and its refill: memory, LDS and
// scalar instructions keep their side of it, VALU instructions may cross (they
// fill the DPP/permlane wait states of the neighbouring step).
constexpr int kSchedValu = 0x2;

// kNoWideStores: no chain kernel stores more than 64 bits per lane in one
// instruction.  A vector store of 96 or 128 bits reads its data VGPRs after
// issue, and hipcc (ROCm 7.2) for gfx950 was seen to place a VALU write of those
// VGPRs (a DPP min of the next step) directly behind such a store with no wait
// state; the D = 256 sweeps then stored partly overwritten data, varying from
// run to run.  With two 64-bit stores (kept apart by a compiler barrier, which
// stops them being merged back) the sweeps are deterministic and bit-identical
// between the column schemes.

// Column sweeps of large volumes keep the backward messages of every
// seg_len()-th node only; the forward pass recomputes a segment's messages from
// the checkpoint at its end, bit for bit (0: every message is stored and
// re-read).
#ifndef CCS_CRF_SEG
#define CCS_CRF_SEG 12
#endif
constexpr int kSeg = CCS_CRF_SEG;
// Segment length by labels per lane: a segment keeps two segments' unaries and
// costs and one segment's messages in registers, so four labels per lane take
// half the length (two waves per SIMD without spills).
template <int LPL>
constexpr int seg_len() {
  return kSeg <= 0 ? 0 : LPL >= 4 ? (kSeg / 2 > 0 ? kSeg / 2 : 1) : kSeg;
}

template <int LPL>
struct Vec {
  float v[LPL];
};

// Row sweeps with a message wave (SEG rows): the backward pass keeps every
// seg_len()-th message in the scratch volume, and a second wave of the
// workgroup recomputes each segment's messages from those checkpoints into an
// LDS ring of kRowRing segments, two segments ahead of the forward pass.  The
// ring holds a node's messages at a stride of 64 * LPL floats.
constexpr int kRowRing = 3;

// Workgroup barrier that orders LDS accesses only: the waves wait for their
// own LDS operations (lgkmcnt) but keep their global loads in flight.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// A node's messages in the LDS ring (lane-major, LPL floats per lane; stores
// of at most 64 bits, as for global memory).
template <int LPL>
__device__ __forceinline__ void st_ring(float* p, int lane, const Vec<LPL>& x) {
  if constexpr (LPL == 4) {
    *reinterpret_cast<float2*>(p + lane * 4) = make_float2(x.v[0], x.v[1]);
    *reinterpret_cast<float2*>(p + lane * 4 + 2) = make_float2(x.v[2], x.v[3]);
  } else if constexpr (LPL == 2) {
    *reinterpret_cast<float2*>(p + lane * 2) = make_float2(x.v[0], x.v[1]);
  } else {
    p[lane] = x.v[0];
  }
}

template <int LPL>
__device__ __forceinline__ Vec<LPL> ld_ring(const float* p, int lane) {
  Vec<LPL> o;
  if constexpr (LPL == 4) {
    const float4 t = *reinterpret_cast<const float4*>(p + lane * 4);
    o.v[0] = t.x; o.v[1] = t.y; o.v[2] = t.z; o.v[3] = t.w;
  } else if constexpr (LPL == 2) {
    const float2 t = *reinterpret_cast<const float2*>(p + lane * 2);
    o.v[0] = t.x; o.v[1] = t.y;
  } else {
    o.v[0] = p[lane];
  }
  return o;
}

// Vector shifted by one label: out(s) = x(s - 1) (dir < 0) or x(s + 1) (dir > 0).
template <int LPL>
__device__ __forceinline__ Vec<LPL> shift_label(const Vec<LPL>& x, int dir, float fill) {
  Vec<LPL> o;
  if (dir < 0) {
    o.v[0] = lane_prev(x.v[LPL - 1], fill);
#pragma unroll
    for (int j = 1; j < LPL; ++j) o.v[j] = x.v[j - 1];
  } else {
    o.v[LPL - 1] = lane_next(x.v[0], fill);
#pragma unroll
    for (int j = 0; j + 1 < LPL; ++j) o.v[j] = x.v[j + 1];
  }
  return o;
}

// The destinations of the one-label lane shifts, carried from step to step.  A
// shift by one lane writes every lane but the boundary one (lane 0 for the
// shift towards higher labels, lane 63 for the other), which keeps the value of
// the destination register; set to the fill once, that lane holds the fill for
// good, and no step re-initialises it.  mn_*: fill +inf (minimum of the
// neighbours), mx_*: fill -inf (maximum).
struct ShiftCarry {
  float mn_prev = kInfF, mn_next = kInfF, mx_prev = -kInfF, mx_next = -kInfF;
};

template <int LPL>
__device__ __forceinline__ Vec<LPL> shift_carry(const Vec<LPL>& x, int dir, float& carry) {
  Vec<LPL> o;
  if (dir < 0) {
    carry = dpp_f<0x138, 0xf, 0xf, false>(carry, x.v[LPL - 1]);  // wave_shr:1
    o.v[0] = carry;
#pragma unroll
    for (int j = 1; j < LPL; ++j) o.v[j] = x.v[j - 1];
  } else {
    carry = dpp_f<0x130, 0xf, 0xf, false>(carry, x.v[0]);  // wave_shl:1
    o.v[LPL - 1] = carry;
#pragma unroll
    for (int j = 0; j + 1 < LPL; ++j) o.v[j] = x.v[j + 1];
  }
  return o;
}

template <int LPL>
__device__ __forceinline__ Vec<LPL> plus(const Vec<LPL>& x, const Vec<LPL>& y) {
  Vec<LPL> o;
#pragma unroll
  for (int j = 0; j < LPL; ++j) o.v[j] = x.v[j] + y.v[j];
  return o;
}

template <int LPL>
__device__ __forceinline__ float local_min(const Vec<LPL>& x) {
  float m = x.v[0];
#pragma unroll
  for (int j = 1; j < LPL; ++j) m = fminf(m, x.v[j]);
  return m;
}

// True in a lane if any of its values is a NaN.  Under -fno-honor-nans the
// compiler may fold any floating-point NaN test (and an integer test of the
// value's bits, which it turns into one) to false, so the bits pass through an
// empty asm first: an integer the compiler knows nothing about.
template <int LPL>
__device__ __forceinline__ bool any_nan(const Vec<LPL>& x) {
  bool r = false;
#pragma unroll
  for (int j = 0; j < LPL; ++j) {
    unsigned u = __float_as_uint(x.v[j]);
    asm volatile("" : "+v"(u));
    r |= (u & 0x7fffffffu) > 0x7f800000u;  // exponent all ones, mantissa not zero
  }
  return r;
}

// The same values in new registers.  The sweeps keep their prefetched nodes in
// a register ring and refill a slot right after reading it; a value read from
// the slot that stays live past the refill would share its register with the
// refill load, and the register allocator would then copy the whole ring at the
// loop back-edge -- a copy that waits for every load in flight.  Copying such
// values first (the copy waits only for this slot's load, issued a ring length
// earlier) keeps each slot in fixed registers.
template <int LPL>
__device__ __forceinline__ Vec<LPL> fresh(const Vec<LPL>& x) {
  Vec<LPL> o;
#pragma unroll
  for (int j = 0; j < LPL; ++j) asm volatile("v_mov_b32 %0, %1" : "=v"(o.v[j]) : "v"(x.v[j]));
  return o;
}

// Makes the values opaque at this point, so that the arithmetic that formed them
// stays ahead of the slot refill that follows (otherwise it may be sunk past the
// refill, which again makes the slot's old value outlive it).
template <int LPL>
__device__ __forceinline__ void pin(Vec<LPL>& x) {
#pragma unroll
  for (int j = 0; j < LPL; ++j) asm volatile("" : "+v"(x.v[j]));
}

// A kernel-uniform value kept in a VGPR.  The steps multiply the edge weight
// (an SGPR, taken with v_readlane) by L1 and L2; with L1 and L2 in VGPRs each of
// those FMAs has one scalar operand, as a VALU instruction allows, and the weight
// needs no copy into a VGPR per step.
__device__ __forceinline__ float in_vgpr(float x) {
  asm volatile("" : "+v"(x));
  return x;
}

// near(s) = min over 0 < |s - t| < delta of x(t)  (+inf if delta == 1).
// DC = 2 fixes delta = 2 at compile time (no branch in the step); DC = 0 reads it.
template <int LPL, int DC>
__device__ __forceinline__ Vec<LPL> near_min(const Vec<LPL>& x, int delta, ShiftCarry& c) {
  if (DC == 2 || delta == 2) {  // the common case, straight-line
    const Vec<LPL> l = shift_carry(x, -1, c.mn_prev), r = shift_carry(x, +1, c.mn_next);
    Vec<LPL> o;
#pragma unroll
    for (int j = 0; j < LPL; ++j) o.v[j] = fminf(l.v[j], r.v[j]);
    return o;
  }
  Vec<LPL> acc;
#pragma unroll
  for (int j = 0; j < LPL; ++j) acc.v[j] = kInfF;
  Vec<LPL> l = x, r = x;
  for (int d = 1; d < delta; ++d) {
    l = shift_label(l, -1, kInfF);
    r = shift_label(r, +1, kInfF);
#pragma unroll
    for (int j = 0; j < LPL; ++j) acc.v[j] = fminf(acc.v[j], fminf(l.v[j], r.v[j]));
  }
  return acc;
}

template <int LPL, int DC>
__device__ __forceinline__ Vec<LPL> near_max(const Vec<LPL>& x, int delta, ShiftCarry& c) {
  if (DC == 2 || delta == 2) {
    const Vec<LPL> l = shift_carry(x, -1, c.mx_prev), r = shift_carry(x, +1, c.mx_next);
    Vec<LPL> o;
#pragma unroll
    for (int j = 0; j < LPL; ++j) o.v[j] = fmaxf(l.v[j], r.v[j]);
    return o;
  }
  Vec<LPL> acc;
#pragma unroll
  for (int j = 0; j < LPL; ++j) acc.v[j] = -kInfF;
  Vec<LPL> l = x, r = x;
  for (int d = 1; d < delta; ++d) {
    l = shift_label(l, -1, -kInfF);
    r = shift_label(r, +1, -kInfF);
#pragma unroll
    for (int j = 0; j < LPL; ++j) acc.v[j] = fmaxf(acc.v[j], fmaxf(l.v[j], r.v[j]));
  }
  return acc;
}

// m(s) = min(U(s), near(U)(s) + P1, min U + P2) with the edge's penalties
// P1 = w L1, P2 = w L2 (PE false: w is the edge weight) or P1 = w, P2 = w2 (PE:
// per-edge penalties).
template <int LPL, int DC, bool PE = false>
__device__ __forceinline__ Vec<LPL> minconv(const Vec<LPL>& U, float umin, float w, float w2, float L1, float L2,
                                            int delta, ShiftCarry& c) {
  const Vec<LPL> nr = near_min<LPL, DC>(U, delta, c);
  const float far = PE ? umin + w2 : umin + w * L2;
  const float wl1 = PE ? w : w * L1;
  Vec<LPL> m;
#pragma unroll
  for (int j = 0; j < LPL; ++j) m.v[j] = fminf(fminf(U.v[j], nr.v[j] + wl1), far);
  return m;
}

// Node loads never sit under divergent control flow (that would make the
// compiler drain every outstanding load at the join): with FULL (D == 64*LPL)
// every lane owns LPL labels and loads them as one vector; otherwise the lane
// loads clamped addresses element by element and masks the labels >= D.
template <int LPL, bool FULL>
__device__ __forceinline__ Vec<LPL> ld_node(const float* __restrict__ base, int lane, int D, float fill) {
  Vec<LPL> out;
  if constexpr (FULL) {
    if constexpr (LPL == 4) {
      const float4 t = *reinterpret_cast<const float4*>(base + lane * 4);
      out.v[0] = t.x; out.v[1] = t.y; out.v[2] = t.z; out.v[3] = t.w;
    } else if constexpr (LPL == 2) {
      const float2 t = *reinterpret_cast<const float2*>(base + lane * 2);
      out.v[0] = t.x; out.v[1] = t.y;
    } else {
      out.v[0] = base[lane];
    }
  } else {
#pragma unroll
    for (int j = 0; j < LPL; ++j) {
      const int s = lane * LPL + j;
      const float v = base[s < D ? s : D - 1];
      out.v[j] = s < D ? v : fill;
    }
  }
  return out;
}

template <int LPL, bool FULL>
__device__ __forceinline__ void st_node(float* __restrict__ base, int lane, int D, const Vec<LPL>& x) {
  if constexpr (FULL) {
    if constexpr (LPL == 4) {  // two 64-bit stores (see kNoWideStores)
      *reinterpret_cast<float2*>(base + lane * 4) = make_float2(x.v[0], x.v[1]);
      asm volatile("" ::: "memory");
      *reinterpret_cast<float2*>(base + lane * 4 + 2) = make_float2(x.v[2], x.v[3]);
    } else if constexpr (LPL == 2) {
      *reinterpret_cast<float2*>(base + lane * 2) = make_float2(x.v[0], x.v[1]);
    } else {
      base[lane] = x.v[0];
    }
  } else {
#pragma unroll
    for (int j = 0; j < LPL; ++j)
      if (lane * LPL + j < D) base[lane * LPL + j] = x.v[j];
  }
}

// Buffer-resource access for the row sweeps.  A row chain's nodes (and its
// message buffer) are contiguous, so a node address is the chain's base, held in
// a buffer resource, plus a 32-bit byte offset in an SGPR: one or two scalar
// instructions per node instead of 64-bit address arithmetic.  Every offset
// passed here is that of a node inside the resource's range.
typedef __amdgpu_buffer_rsrc_t Rsrc;

__device__ __forceinline__ Rsrc chain_rsrc(const float* base, long nfloats) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(base), (short)0, (int)(nfloats * 4), 0x00020000);
}

template <int LPL, bool FULL>
__device__ __forceinline__ Vec<LPL> ld_buf(Rsrc r, int soff, int lane, int D, float fill) {
  Vec<LPL> out;
  if constexpr (FULL && LPL == 4) {
    const u32x4 t = __builtin_amdgcn_raw_buffer_load_b128(r, lane * 16, soff, 0);
#pragma unroll
    for (int j = 0; j < 4; ++j) out.v[j] = __uint_as_float(t[j]);
  } else if constexpr (FULL && LPL == 2) {
    const u32x2 t = __builtin_amdgcn_raw_buffer_load_b64(r, lane * 8, soff, 0);
    out.v[0] = __uint_as_float(t[0]);
    out.v[1] = __uint_as_float(t[1]);
  } else if constexpr (FULL) {
    out.v[0] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, lane * 4, soff, 0));
  } else {
#pragma unroll
    for (int j = 0; j < LPL; ++j) {
      const int s = lane * LPL + j;
      const float v = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, (s < D ? s : D - 1) * 4, soff, 0));
      out.v[j] = s < D ? v : fill;
    }
  }
  return out;
}

template <int LPL, bool FULL>
__device__ __forceinline__ void st_buf(Rsrc r, int soff, int lane, int D, const Vec<LPL>& x) {
  if constexpr (FULL && LPL == 4) {  // two 64-bit stores (see kNoWideStores)
    const u32x2 t = {__float_as_uint(x.v[0]), __float_as_uint(x.v[1])};
    const u32x2 u = {__float_as_uint(x.v[2]), __float_as_uint(x.v[3])};
    __builtin_amdgcn_raw_buffer_store_b64(t, r, lane * 16, soff, 0);
    asm volatile("" ::: "memory");
    __builtin_amdgcn_raw_buffer_store_b64(u, r, lane * 16 + 8, soff, 0);
  } else if constexpr (FULL && LPL == 2) {
    const u32x2 t = {__float_as_uint(x.v[0]), __float_as_uint(x.v[1])};
    __builtin_amdgcn_raw_buffer_store_b64(t, r, lane * 8, soff, 0);
  } else if constexpr (FULL) {
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(x.v[0]), r, lane * 4, soff, 0);
  } else {
#pragma unroll
    for (int j = 0; j < LPL; ++j) {
      const int s = lane * LPL + j;
      if (s < D) __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(x.v[j]), r, s * 4, soff, 0);
    }
  }
}

// Edge weights of a block of up to 64 steps sit one per lane (lane j: the
// block's step j), read from LDS a block ahead; a step takes its weight with
// v_readlane, as an SGPR operand, so the step itself waits on no LDS access.
__device__ __forceinline__ float lane_weight(float wv, int j) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(wv), j));
}

// Suffix sums b_i = u_i + ... + u_{n-2} (b_{n-1} = 0) of the backward-pass
// minima, in float64, the 64 lanes each taking a contiguous segment; returns b_0.
// Done once after the pass so the per-node step carries no float64 work.
__device__ double chain_offsets(const float* u, double* b, int n, int lane) {
  const int m = n - 1;
  const int L = (m + 63) / 64;
  const int s0 = min(lane * L, m), s1 = min(s0 + L, m);
  double tot = 0.0;
  for (int j = s0; j < s1; ++j) tot += (double)u[j];
  double incl = tot;  // sum over this lane and the lanes above it
  for (int o = 1; o < 64; o <<= 1) {
    const double t = __shfl_down(incl, o, 64);
    if (lane + o < 64) incl += t;
  }
  double acc = __shfl_down(incl, 1, 64);
  if (lane == 63) acc = 0.0;
  for (int j = s1 - 1; j >= s0; --j) {
    acc += (double)u[j];
    b[j] = acc;
  }
  if (lane == 0) b[m] = 0.0;
  return __shfl(incl, 0, 64);
}

// Kept constant of the forward pass: r_0 = 0, r_{i+1} = a r_i + v_i - g b_i with
// a = 1 - g for i < m = n - 1; returns g * sum_{i<m} (r_i + b_i) + r_m.  Each
// lane runs its segment from r = 0 and the segments are joined by a scan of the
// affine maps x -> a^len x + r_end.
__device__ double chain_kept(const float* v, const double* b, int n, int lane, double g) {
  const int m = n - 1;
  const double a = 1.0 - g;
  const int L = (m + 63) / 64;
  const int s0 = min(lane * L, m), s1 = min(s0 + L, m);
  double r = 0.0, sr = 0.0, sb = 0.0, geo = 0.0, ap = 1.0;
  for (int j = s0; j < s1; ++j) {
    sr += r;
    geo += ap;
    sb += b[j];
    r = a * r + ((double)v[j] - g * b[j]);
    ap *= a;
  }
  double A = ap, B = r;  // this lane's map, then the composition of lanes 0..lane
  for (int o = 1; o < 64; o <<= 1) {
    const double Ap = __shfl_up(A, o, 64), Bp = __shfl_up(B, o, 64);
    if (lane >= o) {
      B = A * Bp + B;
      A = A * Ap;
    }
  }
  double rin = __shfl_up(B, 1, 64);  // r at the start of this lane's segment
  if (lane == 0) rin = 0.0;
  const double total = wave_sum_f64(g * (sr + rin * geo + sb));
  return total + __shfl(B, 63, 64);
}

// Pipelined chain sweep.  Every input stream of a pass (unaries, cost, backward
// messages, edge weights) runs P nodes ahead of the node being processed: the
// main loop handles blocks of P nodes, unrolled, so ring slot j is a fixed set
// of registers and the load for node s + P is issued right after slot s is read
// (clamped to the last node, so it never leaves the chain).  The last partial
// block is an unrolled tail that issues no loads.
template <int LPL, int MODE, int P, bool FULL, int DC, bool SEG, bool NANCHK, bool PE>
__global__ __launch_bounds__(MODE != kColsMinorant && SEG ? 256 : 64) void k_chain(ChainArgs a) {
  extern __shared__ double lds_all[];
  constexpr bool kCols = MODE == kColsMinorant;
  // Rows with a message wave: wave 0 runs both passes, wave 1 recomputes the
  // forward pass's messages into the LDS ring (see kRowRing).  A workgroup of
  // 256 threads runs two chains (waves 0-1 and 2-3, each with its own LDS), so
  // that between one and two times as many chains as CUs need not put two
  // workgroups on a CU.
  constexpr bool kRowSeg = !kCols && SEG;
  constexpr int KR = kRowSeg ? seg_len<LPL>() : 1;
  static_assert(!kRowSeg || P == kRowRing * KR, "the forward ring spans the message ring");
  static_assert(P <= 64, "one weight per lane and step of a block");
  constexpr int kNS = 64 * LPL;  // ring stride per node, floats
  const int n = a.n, D = a.D;
  const int cs = kRowSeg ? __builtin_amdgcn_readfirstlane(threadIdx.x >> 7) : 0;  // chain slot of the workgroup
  const int cpw = kRowSeg ? (int)(blockDim.x >> 7) : 1;
  const int c_slot = blockIdx.x * cpw + cs;
  const int c = min(c_slot, a.n_chains - 1);  // a missing second chain only keeps the barriers
  const int lane = threadIdx.x & 63;
  const int wid = __builtin_amdgcn_readfirstlane((threadIdx.x >> 6) & (kRowSeg ? 1 : 63));  // wave in the chain
  // per chain: b (doubles), w, u, [w2], v, then the ring (floats, 16-byte aligned)
  const int chain_floats = (((PE ? 6 : 5) * n + 3) & ~3) + (kRowSeg ? kRowRing * KR * kNS : 0);
  double* b_lds = lds_all + cs * (chain_floats / 2);  // b_i = offset of the true backward message at node i
  ShiftCarry cc;
  const float L1v = in_vgpr(a.L1), L2v = in_vgpr(a.L2);
  const long pbase = (long)c * a.pcs;
  const long nstride = a.pns * (long)D;  // floats between consecutive nodes
  // The running state alternates between the two subproblems' unaries, so
  // that each sweep reads one volume for its chain unaries:
  //   rows read A from A_in (A, or the cost volume in a run's first sweep) and
  //     write the column unaries C - A' (A' = A - g);
  //   columns read B = C - A from A and write the row unaries A' = (C - B) + g.
  // Both convert with the cost volume in their forward pass only.  The decode
  // reads A.
  const float* __restrict__ Ab = (kCols ? a.A : a.A_in) + pbase * D;
  float* __restrict__ Aw = a.A + pbase * D;
  const float* __restrict__ Cb = MODE != kRowsDecode ? a.cost + pbase * D : nullptr;
  const float* __restrict__ wb = a.w + pbase * a.wmul;
  const float* __restrict__ wb2 = PE ? a.w2 + pbase * a.wmul : nullptr;
  float* __restrict__ msg = a.msg + (long)c * n * D;
  bool valid[LPL];
#pragma unroll
  for (int j = 0; j < LPL; ++j) valid[j] = FULL || lane * LPL + j < D;

  // The chain unaries: what the sweep reads (A for rows, B for columns).
  auto theta_of = [&](const Vec<LPL>& av) {
    Vec<LPL> th;
#pragma unroll
    for (int j = 0; j < LPL; ++j) th.v[j] = valid[j] ? av.v[j] : kInfF;
    return th;
  };
  // Columns: the row unaries C - B of a node.
  auto row_unaries = [&](const Vec<LPL>& bv, const Vec<LPL>& cv) {
    Vec<LPL> r;
#pragma unroll
    for (int j = 0; j < LPL; ++j) r.v[j] = cv.v[j] - bv.v[j];
    return r;
  };
  // Rows: the column unaries C - A' stored for a node.
  auto col_unaries = [&](const Vec<LPL>& out, const Vec<LPL>& cv) {
    Vec<LPL> r;
#pragma unroll
    for (int j = 0; j < LPL; ++j) r.v[j] = cv.v[j] - out.v[j];
    return r;
  };

  // The chain's edge weights (per-edge P1 and P2 with PE) go to LDS once (the
  // sweeps read them a block of steps at a time, one per lane).  u_lds takes
  // the per-node minima of a pass (float64 sums are formed after it).
  float* w_lds = reinterpret_cast<float*>(b_lds + n);
  float* u_lds = w_lds + n;
  float* w2_lds = u_lds + n;  // PE only
  // kRowSeg: the forward minima go to v_lds (u_lds keeps the backward minima
  // that the message wave reads), then the message ring at a 16-byte boundary.
  float* v_lds = w2_lds + (PE ? n : 0);
  // (offset arithmetic on the LDS pointer itself keeps its address space)
  float* ring = reinterpret_cast<float*>(b_lds) + (((PE ? 6 : 5) * n + 3) & ~3);
  float* fw_lds = kRowSeg ? v_lds : u_lds;
  const long wstep = a.pns * a.wmul;
  constexpr int kChainThreads = kRowSeg ? 128 : 64;
  for (int i = threadIdx.x & (kChainThreads - 1); i < n - 1; i += kChainThreads) {
    w_lds[i] = wb[(long)i * wstep];
    if constexpr (PE) w2_lds[i] = wb2[(long)i * wstep];
  }
  __syncthreads();
  if constexpr (kRowSeg) {
    if (c_slot >= a.n_chains) {  // the workgroup's second chain is missing: the same barriers, no work
      const int M = (n - 1 + KR - 1) / KR;
      __syncthreads();
      for (int k = 0; k < (M == 0 ? 0 : (M > 2 ? M - 1 : 1)); ++k) lds_barrier();
      return;
    }
  }

  Vec<LPL> ra[P], rc[P], rb[P];
  // Row chains (the host guarantees a row's byte extent fits in 31 bits): the
  // unaries and the messages through buffer resources; columns use 64-bit
  // addresses (their nodes span the whole volume).
  const int strideB = (int)(nstride * 4);
  const Rsrc rA = chain_rsrc(Ab, (long)(n - 1) * nstride + D);
  const Rsrc rW = chain_rsrc(Aw, (long)(n - 1) * nstride + D);
  const Rsrc rM = chain_rsrc(msg, (long)n * D);

  // ---- backward pass: step s handles i = n-2-s: Bh[i] = minconv(theta[i+1] + Bh[i+1]) - min(...)
  // kRowSeg: wave 0 only; it stores the messages at i = nb - m*KR (m >= 1), so
  // a step s0 + j of a block (s0 a multiple of P, itself a multiple of KR) is a
  // checkpoint exactly when (j + 1) % KR == 0.
  const int nb = n - 1;
  Vec<LPL> bh;
  bool nan_seen = false;  // NANCHK: a unary read by this lane was NaN
  double bsum = 0.0;
  Vec<LPL> a0, c0, th0;
  // kRowSeg, wave 1: the forward pass's messages.  Segment s holds the forward
  // steps i' = s*KR + t (t < KR) of node i = i' - g, where the g = M*KR - nb
  // virtual steps before node 0 end-align the segments with the checkpoints
  // (Bh at (s+1)*KR - g is one; Bh_{n-1} = 0).  Each message is recomputed
  // with the backward step's own operations and operands (theta_{i+1} + Bh_{i+1},
  // the stored minimum u_i, the edge's penalties), so it equals the message
  // the backward pass formed, and goes to ring slot i' % (kRowRing*KR).
  // Segments 0 and 1 are done before the first barrier, then segment s before
  // the barrier that wave 0 passes at the start of segment s - 1.
  auto row_messages = [&]() {
    if constexpr (kRowSeg) {
      constexpr int K = KR;
      const int M = (nb + K - 1) / K;
      // NANCHK: this wave tests node 0 and every unary its segments consume
      // (nodes 1 .. n-1) and sets the chain's flag, off the forward wave.
      bool nan1 = NANCHK && any_nan(ld_buf<LPL, FULL>(rA, 0, lane, D, kInfF));
      auto put_flag = [&]() {
        if constexpr (NANCHK) {
          const bool any = __ballot(nan1) != 0;  // all lanes take part
          if (lane == 0) a.nan_flags[c] = any ? 1 : 0;
        }
      };
      if (M == 0) {
        put_flag();
        return;
      }
      const int g = M * K - nb;
      struct HSeg {
        Vec<LPL> a[K], m;
        float u, w, w2;
      };
      HSeg X, Y;
      auto load_h = [&](int s, HSeg& B) {  // s is clamped by the caller to a segment
        const int ib = s * K - g;            // node of t = 0 (negative only in segment 0)
#pragma unroll
        for (int t = 0; t < K; ++t)
          B.a[t] = ld_buf<LPL, FULL>(rA, min(max(ib + t + 1, 1), n - 1) * strideB, lane, D, kInfF);
        B.m = ld_buf<LPL, FULL>(rM, (ib + K) * D * 4, lane, D, 0.f);
        const int e = max(min(ib + lane, n - 2), 0);
        B.u = u_lds[e];
        B.w = w_lds[e];
        B.w2 = PE ? w2_lds[e] : 0.f;
      };
      auto run_h = [&](int s, const HSeg& B) {
        const int ib = s * K - g;
        float* slot = ring + (s % kRowRing) * K * kNS;
        Vec<LPL> Xb = B.m;
#pragma unroll
        for (int t = K - 1; t >= 0; --t) {
          if (ib + t >= 0) {
            if constexpr (NANCHK) nan1 |= any_nan(B.a[t]);
            const Vec<LPL> U = plus(theta_of(B.a[t]), Xb);
            const float u = lane_weight(B.u, t);
            const Vec<LPL> m = minconv<LPL, DC, PE>(U, u, lane_weight(B.w, t), PE ? lane_weight(B.w2, t) : 0.f,
                                                    L1v, L2v, a.delta, cc);
#pragma unroll
            for (int k = 0; k < LPL; ++k) Xb.v[k] = valid[k] ? m.v[k] - u : 0.f;
            st_ring<LPL>(slot + t * kNS, lane, Xb);
          }
        }
      };
      load_h(0, X);
      load_h(min(1, M - 1), Y);
      run_h(0, X);
      load_h(min(2, M - 1), X);
      if (M > 1) run_h(1, Y);
      lds_barrier();  // segments 0 and 1 are in the ring
      for (int s = 2; s < M;) {
        load_h(min(s + 1, M - 1), Y);
        __builtin_amdgcn_sched_barrier(0);
        run_h(s, X);
        lds_barrier();
        if (++s >= M) break;
        load_h(min(s + 1, M - 1), X);
        __builtin_amdgcn_sched_barrier(0);
        run_h(s, Y);
        lds_barrier();
        ++s;
      }
      put_flag();
    }
  };
  if (!kRowSeg || wid == 0) {
  auto load_b = [&](int s, Vec<LPL>& av) {
    const int node = s < nb ? n - 1 - s : 1;  // clamped: beyond the chain re-read node 1
    if constexpr (kCols)
      av = ld_node<LPL, FULL>(Ab + node * nstride, lane, D, kInfF);
    else
      av = ld_buf<LPL, FULL>(rA, node * strideB, lane, D, kInfF);
  };
  // weights of steps s0 .. s0+63 of the backward pass (edge n-2-s)
  auto wlane_b = [&](int s0) { return w_lds[max(n - 2 - (s0 + lane), 0)]; };
  auto wlane_b2 = [&](int s0) { return PE ? w2_lds[max(n - 2 - (s0 + lane), 0)] : 0.f; };
#pragma unroll
  for (int j = 0; j < LPL; ++j) bh.v[j] = 0.f;
  st_node<LPL, FULL>(msg + (long)(n - 1) * D, lane, D, bh);
  auto plus_bh = [&](const Vec<LPL>& th) {
    Vec<LPL> U;
#pragma unroll
    for (int k = 0; k < LPL; ++k) U.v[k] = th.v[k] + bh.v[k];
    return U;
  };
  // ck: (kRowSeg) this step's message is a checkpoint
  auto step_b = [&](int s, const Vec<LPL>& U, float wi, float wi2, bool ck) {
    const int i = n - 2 - s;
    const float u = wave_min(local_min(U));
    const Vec<LPL> m = minconv<LPL, DC, PE>(U, u, wi, wi2, L1v, L2v, a.delta, cc);
#pragma unroll
    for (int k = 0; k < LPL; ++k) bh.v[k] = valid[k] ? m.v[k] - u : 0.f;
    if constexpr (kCols) {
      constexpr int K = SEG ? seg_len<LPL>() : 0;
      if (K == 0 || i % K == 0) st_node<LPL, FULL>(msg + (long)i * D, lane, D, bh);  // i is wave-uniform
    } else if constexpr (kRowSeg) {
      if (ck) st_buf<LPL, FULL>(rM, i * D * 4, lane, D, bh);
    } else {
      st_buf<LPL, FULL>(rM, i * D * 4, lane, D, bh);
    }
    u_lds[i] = u;  // every lane writes the same value
  };
  if (nb > 0) {
#pragma unroll
    for (int j = 0; j < P; ++j) load_b(j, ra[j]);
  }
  int s0 = 0;
  float wv = wlane_b(0), wv2 = wlane_b2(0);
  for (; s0 + P <= nb; s0 += P) {
    const float wnext = wlane_b(s0 + P), wnext2 = wlane_b2(s0 + P);
    // kRowSeg: byte offsets of the block's refills, one per lane (v_readlane
    // per step instead of clamping each offset with scalar instructions)
    const int bo = kRowSeg ? (s0 + P + lane < nb ? n - 1 - (s0 + P + lane) : 1) * strideB : 0;
#pragma unroll
    for (int j = 0; j < P; ++j) {
      if constexpr (NANCHK && !kRowSeg) nan_seen |= any_nan(ra[j]);
      // Slot j is consumed (U is a new value) before it is refilled, so the
      // refill reuses the slot's registers.
      const Vec<LPL> U = plus_bh(theta_of(ra[j]));
      __builtin_amdgcn_sched_barrier(kSchedValu);
      if constexpr (kRowSeg)
        ra[j] = ld_buf<LPL, FULL>(rA, __builtin_amdgcn_readlane(bo, j), lane, D, kInfF);
      else
        load_b(s0 + j + P, ra[j]);
      step_b(s0 + j, U, lane_weight(wv, j), PE ? lane_weight(wv2, j) : 0.f, (j + 1) % KR == 0);
    }
    wv = wnext;
    wv2 = wnext2;
  }
#pragma unroll
  for (int j = 0; j < P; ++j)
    if (s0 + j < nb) {
      if constexpr (NANCHK && !kRowSeg) nan_seen |= any_nan(ra[j]);
      step_b(s0 + j, plus_bh(theta_of(ra[j])), lane_weight(wv, j), PE ? lane_weight(wv2, j) : 0.f,
             (j + 1) % KR == 0);
    }
  bsum = chain_offsets(u_lds, b_lds, n, lane);  // b_0 = sum of all minima

  // bh holds Bh[0]; theta[0] is read directly.
  a0 = ld_node<LPL, FULL>(Ab, lane, D, kInfF);
  if constexpr (NANCHK && !kRowSeg) {  // (kRowSeg: the message wave tests the unaries)
    nan_seen |= any_nan(a0);
    const bool any = __ballot(nan_seen) != 0;  // all lanes take part
    if (lane == 0) a.nan_flags[c] = any ? 1 : 0;
  }
  if constexpr (MODE != kRowsDecode) c0 = ld_node<LPL, FULL>(Cb, lane, D, kInfF);
  th0 = theta_of(a0);
  }
  __syncthreads();  // kRowSeg: the checkpoints and u_lds are now visible to wave 1
  if constexpr (kRowSeg) {
    if (wid == 1) {
      row_messages();
      return;
    }
  }
  Vec<LPL> M0;
#pragma unroll
  for (int k = 0; k < LPL; ++k) M0.v[k] = th0.v[k] + bh.v[k];
  const double Lambda = bsum + (double)wave_min(local_min(M0));

  {
    // Forward pass: keep g_i = gamma (R + Bh) + excess, pass V = R - gamma (R + Bh) on.
    // Step i's operands: Bh_i, the read unaries of node i+1 and the costs of
    // node i+1 (columns: they form A_{i+1}) or of node i (rows: they form the
    // stored C - A'_i).
    const Rsrc rC = chain_rsrc(Cb, kCols ? 0 : (long)(n - 1) * nstride + D);  // (rows)
    auto load_f = [&](int i, Vec<LPL>& bv, Vec<LPL>& av, Vec<LPL>& cv) {
      const int node = i < nb ? i : nb - 1;
      if constexpr (kCols) {
        bv = ld_node<LPL, FULL>(msg + (long)node * D, lane, D, 0.f);
        av = ld_node<LPL, FULL>(Ab + (node + 1) * nstride, lane, D, kInfF);
        cv = ld_node<LPL, FULL>(Cb + (node + 1) * nstride, lane, D, kInfF);
      } else {
        bv = ld_buf<LPL, FULL>(rM, node * D * 4, lane, D, 0.f);
        av = ld_buf<LPL, FULL>(rA, (node + 1) * strideB, lane, D, kInfF);
        cv = ld_buf<LPL, FULL>(rC, node * strideB, lane, D, kInfF);
      }
    };
    // weights of steps i0 .. i0+63 of the forward pass (edge i)
    auto wlane_f = [&](int i0) { return w_lds[max(min(i0 + lane, n - 2), 0)]; };
    auto wlane_f2 = [&](int i0) { return PE ? w2_lds[max(min(i0 + lane, n - 2), 0)] : 0.f; };
    Vec<LPL> R = th0;
    Vec<LPL> a_cur = kCols ? row_unaries(a0, c0) : a0;  // A_0
    // Rows: the costs of the last node, for its stored C - A'.
    Vec<LPL> c_last;
    if constexpr (!kCols) c_last = ld_buf<LPL, FULL>(rC, (n - 1) * strideB, lane, D, kInfF);
    // The float64 kept constant (the chain minimum this pass leaves behind) is
    // formed from the per-node minima v_i after the pass: chain_kept().
    // step_f takes RB = R + Bh_i, A_{i+1} and theta_{i+1} (the read unaries).
    // step_f returns the node's new unaries; the caller stores them.
    auto step_f = [&](int i, const Vec<LPL>& RB, const Vec<LPL>& an, const Vec<LPL>& thn, float wi, float wi2) {
      Vec<LPL> gt, V;
#pragma unroll
      for (int k = 0; k < LPL; ++k) {
        gt.v[k] = valid[k] ? kGamma * RB.v[k] : 0.f;
        V.v[k] = valid[k] ? R.v[k] - gt.v[k] : kInfF;
      }
      const float v = wave_min(local_min(V));
      fw_lds[i] = v;  // every lane writes the same value
      const Vec<LPL> m = minconv<LPL, DC, PE>(V, v, wi, wi2, L1v, L2v, a.delta, cc);
      Vec<LPL> mm;
#pragma unroll
      for (int k = 0; k < LPL; ++k) mm.v[k] = valid[k] ? m.v[k] : -kInfF;
      const Vec<LPL> nmax = near_max<LPL, DC>(mm, a.delta, cc);
      const float wl1 = PE ? wi : wi * L1v;
      Vec<LPL> out;
#pragma unroll
      for (int k = 0; k < LPL; ++k) {
        // g = gamma (R + Bh) + excess, excess = V - max(m, near_max(m) - w L1); with
        // V = R - gamma (R + Bh) that is R - max(...), formed directly.
        const float g = valid[k] ? R.v[k] - fmaxf(m.v[k], nmax.v[k] - wl1) : 0.f;
        out.v[k] = MODE == kRowsMinorant ? a_cur.v[k] - g : a_cur.v[k] + g;
        R.v[k] = valid[k] ? thn.v[k] + m.v[k] - v : kInfF;
      }
      a_cur = an;
      return out;
    };
    // Stores node i's result: columns A'_i, rows C_i - A'_i (ci: C_i).
    auto store_f = [&](int i, const Vec<LPL>& out, const Vec<LPL>& ci) {
      if constexpr (kCols) st_node<LPL, FULL>(Aw + i * nstride, lane, D, out);
      else st_buf<LPL, FULL>(rW, i * strideB, lane, D, col_unaries(out, ci));
    };
    if constexpr (kRowSeg) {
      // Forward step i' = i + g (see row_messages): A_{i+1} in register ring
      // slot i' % P (P = kRowRing*KR, so the segments i' = q*KR .. q*KR+KR-1
      // line up with the blocks), Bh_i from LDS ring slot i' % P, read two
      // steps ahead.  Segments 0 and 1 are complete at the first barrier;
      // before the first step of segment q (1 <= q <= M-2) a barrier pairs
      // with wave 1's barrier after segment q+1, which also frees the slots of
      // segment q-1 for segment q+2.
      constexpr int K = KR;
      const int M = (nb + K - 1) / K;
      const int g = M * K - nb;
      const int nbp = M * K;
      // Register ring slot j (steps ip = j mod P) holds A_{i+1}, read by step
      // ip as theta_{i+1} and by step ip + 1 as its own node's unaries A_i; it
      // is refilled after that second read, so neither read copies it out.
      // Node of step ip's A (A_{i+1}, i = ip - g); virtual steps (i < 0) read
      // A_0, which the first real step reads as its own node's unaries.
      auto node_a = [&](int ip) { return max(min(ip - g, nb - 1) + 1, 0); };
      // rc slot j holds C_i of the node step ip stores, refilled after the store.
      auto node_c = [&](int ip) { return max(min(ip - g, nb - 1), 0); };
      // Byte offsets of a block's refills (A of step ip0 + P + lane), one per
      // lane: a step takes its offset with v_readlane instead of clamping it.
      auto refill_offsets = [&](int ip0) { return node_a(ip0 + P + lane) * strideB; };
      auto wl = [&](int ip0) { return w_lds[max(min(ip0 - g + lane, n - 2), 0)]; };
      auto wl2 = [&](int ip0) { return PE ? w2_lds[max(min(ip0 - g + lane, n - 2), 0)] : 0.f; };
      Vec<LPL> bm[3];
      if (nb > 0) {
        // slots 0 .. P-2: steps 0 .. P-2; slot P-1: A of step -1 (A_0, the first
        // step's own unaries), its C that of step P-1
#pragma unroll
        for (int j = 0; j < P; ++j) {
          ra[j] = ld_buf<LPL, FULL>(rA, node_a(j < P - 1 ? j : -1) * strideB, lane, D, kInfF);
          rc[j] = ld_buf<LPL, FULL>(rC, node_c(j) * strideB, lane, D, kInfF);
        }
        lds_barrier();  // segments 0 and 1 are in the ring
        bm[0] = ld_ring<LPL>(ring, lane);
        bm[1] = ld_ring<LPL>(ring + kNS, lane);
      }
      float wf = wl(0), wf2 = wl2(0);
      // Step j of the block at ip0 (ao: the block's refill offsets, aop: the
      // previous block's).  LOW: the block may start with virtual steps (i < 0,
      // skipped); TAIL: no refills, steps past the chain skipped.
      auto fstep = [&](int ip0, int ao, int aop, int j, bool LOW, bool TAIL) {
        const int ip = ip0 + j;
        if (TAIL && ip >= nbp) return;
        if (j % K == 0) {
          const int q = ip / K;
          if (q >= 1 && q <= M - 2) lds_barrier();
        }
        const int i = ip - g;
        const int jp = (j + P - 1) % P;  // the previous step's slot
        a_cur = ra[jp];                  // A_i
        const Vec<LPL> RB = plus(R, bm[j % 3]);
        __builtin_amdgcn_sched_barrier(kSchedValu);
        bm[(j + 2) % 3] = ld_ring<LPL>(ring + ((j + 2) % P) * kNS, lane);
        if (!LOW || i >= 0)
          store_f(i, step_f(i, RB, ra[j], theta_of(ra[j]), lane_weight(wf, j), PE ? lane_weight(wf2, j) : 0.f),
                  rc[j]);
        if (!TAIL) {
          // C_i and A_i are read: their slots take C of step ip + P and A of
          // step ip - 1 + P.
          rc[j] = ld_buf<LPL, FULL>(rC, __builtin_amdgcn_readlane(ao, j) - strideB, lane, D, kInfF);
          ra[jp] = ld_buf<LPL, FULL>(rA, j > 0 ? __builtin_amdgcn_readlane(ao, j - 1) : __builtin_amdgcn_readlane(aop, P - 1),
                                     lane, D, kInfF);
        }
      };
      int i0p = 0;
      if (nbp >= P) {
        int aop = refill_offsets(-P);
        {
          const float wn = wl(P), wn2 = wl2(P);
          const int ao = refill_offsets(0);
#pragma unroll
          for (int j = 0; j < P; ++j) fstep(0, ao, aop, j, true, false);
          aop = ao;
          wf = wn;
          wf2 = wn2;
        }
        for (i0p = P; i0p + P <= nbp; i0p += P) {
          const float wn = wl(i0p + P), wn2 = wl2(i0p + P);
          const int ao = refill_offsets(i0p);
#pragma unroll
          for (int j = 0; j < P; ++j) fstep(i0p, ao, aop, j, false, false);
          aop = ao;
          wf = wn;
          wf2 = wn2;
        }
      }
      // (the last node's own unaries: a_cur as the last step left it)
#pragma unroll
      for (int j = 0; j < P; ++j) fstep(i0p, 0, 0, j, true, true);
    } else
    if constexpr (kCols && SEG && seg_len<LPL>() > 0) {
      // Segments of K forward steps i0 .. i0+K-1.  Their nodes i0+1 .. i0+K
      // (unaries and costs), the checkpoint Bh_{min(i0+K, n-1)} (Bh_{n-1} = 0
      // is stored) and the segment's minima u_i and weights w_i (one per lane)
      // are loaded a segment ahead.  Bh_i for the segment is recomputed
      // backwards from the checkpoint with the backward step's own operations
      // and operands (theta_{i+1}, u_i, w_i), so it equals the message the
      // backward pass formed; u_lds[i] still holds the backward minimum until
      // forward step i overwrites it.
      constexpr int K = seg_len<LPL>();
      struct Seg {
        Vec<LPL> a[K], c[K], m;
        float u, w, w2;
      };
      Seg X, Y;
      Vec<LPL> bs[K];
      // Nodes first .. first+K-1 (clamped to the chain) through buffer
      // resources based at the segment: 32-bit offsets, one scalar multiply
      // per node (the host bounds K nodes of a column by 2 GiB).
      auto load_seg = [&](int i0, Seg& B) {
        const int first = min(i0 + 1, n - 1);
        const long span = (long)(K - 1) * nstride + D;
        const Rsrc sA = chain_rsrc(Ab + first * nstride, span);
        const Rsrc sC = chain_rsrc(Cb + first * nstride, span);
#pragma unroll
        for (int t = 0; t < K; ++t) {
          const int off = min(t, n - 1 - first) * strideB;
          B.a[t] = ld_buf<LPL, FULL>(sA, off, lane, D, kInfF);
          B.c[t] = ld_buf<LPL, FULL>(sC, off, lane, D, kInfF);
        }
        B.m = ld_node<LPL, FULL>(msg + (long)min(i0 + K, nb) * D, lane, D, 0.f);
        const int e = max(min(i0 + lane, n - 2), 0);
        B.u = u_lds[e];
        B.w = w_lds[e];
        B.w2 = PE ? w2_lds[e] : 0.f;
      };
      // The segment of cnt (1 .. K) steps from i0, its data in B.
      auto run_seg = [&](int i0, int cnt, const Seg& B) {
        Vec<LPL> Xb = B.m;
#pragma unroll
        for (int t = K - 1; t >= 0; --t) {
          if (t < cnt) {
            const Vec<LPL> U = plus(theta_of(B.a[t]), Xb);
            const float u = lane_weight(B.u, t);
            const Vec<LPL> m = minconv<LPL, DC, PE>(U, u, lane_weight(B.w, t), PE ? lane_weight(B.w2, t) : 0.f,
                                                    L1v, L2v, a.delta, cc);
#pragma unroll
            for (int k = 0; k < LPL; ++k) Xb.v[k] = valid[k] ? m.v[k] - u : 0.f;
            bs[t] = Xb;
          }
        }
        const Rsrc sW = chain_rsrc(Aw + i0 * nstride, (long)(K - 1) * nstride + D);
#pragma unroll
        for (int t = 0; t < K; ++t)
          if (t < cnt)
            st_buf<LPL, FULL>(sW, t * strideB, lane, D,
                              step_f(i0 + t, plus(R, bs[t]), row_unaries(B.a[t], B.c[t]), theta_of(B.a[t]),
                                     lane_weight(B.w, t), PE ? lane_weight(B.w2, t) : 0.f));
      };
      // Two segments per iteration, alternating between X and Y; the next
      // segment's loads are issued before the current segment's steps.
      load_seg(0, X);
      int i0 = 0;
      for (;;) {
        if (i0 + K >= nb) break;  // X holds the last segment
        load_seg(i0 + K, Y);
        __builtin_amdgcn_sched_barrier(0);
        run_seg(i0, K, X);
        i0 += K;
        if (i0 + K >= nb) {
          X = Y;
          break;
        }
        load_seg(i0 + K, X);
        __builtin_amdgcn_sched_barrier(0);
        run_seg(i0, K, Y);
        i0 += K;
      }
      if (i0 < nb) run_seg(i0, nb - i0, X);
    } else {
    if (nb > 0) {
#pragma unroll
      for (int j = 0; j < P; ++j) load_f(j, rb[j], ra[j], rc[j]);
    }
    int i0 = 0;
    float wf = wlane_f(0), wf2 = wlane_f2(0);
    for (; i0 + P <= nb; i0 += P) {
      const float wnext = wlane_f(i0 + P), wnext2 = wlane_f2(i0 + P);
#pragma unroll
      for (int j = 0; j < P; ++j) {
        // A_{i+1} lives on as the next step's a_cur: copied out of the slot
        // (columns form it from C and B).
        Vec<LPL> an, thn;
        if constexpr (kCols) {
          an = row_unaries(ra[j], rc[j]);
          thn = theta_of(ra[j]);
          pin(an);
          pin(thn);
        } else {
          an = fresh(ra[j]);
          thn = an;
        }
        const Vec<LPL> ci = fresh(rc[j]);  // (rows: C_i)
        Vec<LPL> RB = plus(R, rb[j]);
        if constexpr (kCols) pin(RB);
        __builtin_amdgcn_sched_barrier(kSchedValu);
        load_f(i0 + j + P, rb[j], ra[j], rc[j]);
        store_f(i0 + j, step_f(i0 + j, RB, an, thn, lane_weight(wf, j), PE ? lane_weight(wf2, j) : 0.f), ci);
      }
      wf = wnext;
      wf2 = wnext2;
    }
#pragma unroll
    for (int j = 0; j < P; ++j)
      if (i0 + j < nb)
        store_f(i0 + j,
                step_f(i0 + j, plus(R, rb[j]), kCols ? row_unaries(ra[j], rc[j]) : ra[j], theta_of(ra[j]),
                       lane_weight(wf, j), PE ? lane_weight(wf2, j) : 0.f),
                rc[j]);
    }
    {
      Vec<LPL> out;
#pragma unroll
      for (int k = 0; k < LPL; ++k)
        out.v[k] = MODE == kRowsMinorant ? a_cur.v[k] - R.v[k] : a_cur.v[k] + R.v[k];
      if constexpr (kCols)
        st_node<LPL, FULL>(Aw + (n - 1) * nstride, lane, D, out);
      else
        st_node<LPL, FULL>(Aw + (n - 1) * nstride, lane, D, col_unaries(out, c_last));
    }
    const double kept = chain_kept(fw_lds, b_lds, n, lane, kGammaD);
    if (lane == 0) {
      a.lam[c] = Lambda;
      a.kept[c] = kept;
    }
  }
}

// Sum of u_0 .. u_{n-2} in float64 (64 lanes, contiguous segments).
__device__ double chain_total(const float* u, int n, int lane) {
  const int m = n - 1;
  const int L = (m + 63) / 64;
  const int s0 = min(lane * L, m), s1 = min(s0 + L, m);
  double tot = 0.0;
  for (int j = s0; j < s1; ++j) tot += (double)u[j];
  return wave_sum_f64(tot);
}

// Runs step(i, A_{i+aoff}, msg_i, w_{i+woff}) for the cnt nodes i = first,
// first + dir, ..., with the unaries (and, with MSG, the message buffer) P nodes
// ahead in a register ring, as in k_chain, and the edge weights w (in LDS, n - 1
// of them) one block ahead in lanes.
template <int LPL, bool FULL, int P, bool MSG, bool PE, typename Step>
__device__ __forceinline__ void ring_pass(Rsrc rA, int strideB, Rsrc rM, const float* w_lds, const float* w2_lds,
                                          int n, int D, int lane, int first, int cnt, int dir, int aoff, int woff,
                                          Step&& step) {
  static_assert(P <= 64, "one weight per lane and step of a block");
  if (cnt <= 0) return;
  Vec<LPL> ra[P], rm[P];
  auto load = [&](int s, Vec<LPL>& av, Vec<LPL>& mv) {
    const int i = first + dir * (s < cnt ? s : cnt - 1);
    av = ld_buf<LPL, FULL>(rA, (i + aoff) * strideB, lane, D, kInfF);
    if constexpr (MSG) mv = ld_buf<LPL, FULL>(rM, i * D * 4, lane, D, 0.f);
  };
  auto wlane = [&](int s0) { return w_lds[max(min(first + dir * (s0 + lane) + woff, n - 2), 0)]; };
  auto wlane2 = [&](int s0) { return PE ? w2_lds[max(min(first + dir * (s0 + lane) + woff, n - 2), 0)] : 0.f; };
#pragma unroll
  for (int j = 0; j < P; ++j) load(j, ra[j], rm[j]);
  int s0 = 0;
  float wv = wlane(0), wv2 = wlane2(0);
  for (; s0 + P <= cnt; s0 += P) {
    const float wnext = wlane(s0 + P), wnext2 = wlane2(s0 + P);
#pragma unroll
    for (int j = 0; j < P; ++j) {
      const Vec<LPL> av = fresh(ra[j]);
      Vec<LPL> mv;
      if constexpr (MSG) mv = fresh(rm[j]);
      __builtin_amdgcn_sched_barrier(kSchedValu);
      load(s0 + j + P, ra[j], rm[j]);
      step(first + dir * (s0 + j), av, mv, lane_weight(wv, j), PE ? lane_weight(wv2, j) : 0.f);
    }
    wv = wnext;
    wv2 = wnext2;
  }
#pragma unroll
  for (int j = 0; j < P; ++j)
    if (s0 + j < cnt) step(first + dir * (s0 + j), ra[j], rm[j], lane_weight(wv, j), PE ? lane_weight(wv2, j) : 0.f);
}

// Row decode on four waves.  Wave 0 runs the forward messages F from node 0,
// wave 1 the backward messages Bh from node n-1.  Until the midpoint each
// stores its messages in the scratch volume (phase 1); then each goes on into
// the other half (phase 2) and hands its messages to a decision wave through
// an LDS ring of two segments of K nodes: wave 2 decides the nodes mid .. n-1
// (F_i from wave 0's ring, Bh_i from the scratch volume), wave 3 the nodes
// mid-1 .. 0 (Bh_i from wave 1's ring, F_i from the scratch volume).  Phase 2
// runs in rounds: in round r the message waves write segment r into slot r % 2
// and the decision waves decide segment r - 1 from slot (r - 1) % 2; all four
// waves meet at a barrier after every round.  F_i + Bh_i and the decision rule
// are those of the single-wave decode, so the labels are the same; the chain
// latency is n message steps, with no decision work on the message waves.
template <int LPL, int P, bool FULL, int DC, bool GAP, bool PE>
__global__ __launch_bounds__(256) void k_decode4(ChainArgs a) {
  extern __shared__ double lds2[];
  constexpr int K = seg_len<LPL>();
  static_assert(P == kRowRing * K, "the message waves' register ring spans three segments");
  static_assert(P <= 64, "one weight per lane and step of a block");
  constexpr int kNS = 64 * LPL;  // ring stride per node, floats
  const int c = blockIdx.x;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform branches
  const int lane = threadIdx.x & 63;
  const int n = a.n, D = a.D;
  ShiftCarry cc;
  const float L1v = in_vgpr(a.L1), L2v = in_vgpr(a.L2);
  const long pbase = (long)c * a.pcs;
  const long nstride = a.pns * (long)D;
  const float* __restrict__ Ab = a.A + pbase * D;
  const float* __restrict__ wb = a.w + pbase * a.wmul;
  const float* __restrict__ wb2 = PE ? a.w2 + pbase * a.wmul : nullptr;
  float* msg = a.msg + (long)c * n * D;
  float* w_lds = reinterpret_cast<float*>(lds2);
  float* u_lds = w_lds + n;
  float* w2_lds = u_lds + n;  // PE only
  float* ring = w_lds + (((PE ? 3 : 2) * n + 3) & ~3);  // [half][slot][K][kNS]
  const int strideB = (int)(nstride * 4);
  const Rsrc rA = chain_rsrc(Ab, (long)(n - 1) * nstride + D);
  const Rsrc rM = chain_rsrc(msg, (long)n * D);
  bool valid[LPL];
#pragma unroll
  for (int j = 0; j < LPL; ++j) valid[j] = FULL || lane * LPL + j < D;
  const long wstep = a.pns * a.wmul;
  for (int i = threadIdx.x; i < n - 1; i += 256) {
    w_lds[i] = wb[(long)i * wstep];
    if constexpr (PE) w2_lds[i] = wb2[(long)i * wstep];
  }
  __syncthreads();
  const int mid = n / 2;
  const int len0 = n - mid, len1 = mid;  // phase-2 nodes of waves 0 / 1
  const int S = ((len0 > len1 ? len0 : len1) + K - 1) / K;  // rounds with a segment

  Vec<LPL> X;  // wave 0: F_i, wave 1: Bh_i
  auto fwd = [&](const Vec<LPL>& an, float w, float w2) {  // F_i = A_i + minconv(F_{i-1}; w_{i-1}) - min F_{i-1}
    const float f = wave_min(local_min(X));
    const Vec<LPL> m = minconv<LPL, DC, PE>(X, f, w, w2, L1v, L2v, a.delta, cc);
#pragma unroll
    for (int k = 0; k < LPL; ++k) X.v[k] = valid[k] ? an.v[k] + m.v[k] - f : kInfF;
  };
  auto bwd = [&](int i, const Vec<LPL>& an1, float w, float w2) {  // Bh_i from Bh_{i+1}, A_{i+1} and w_i
    Vec<LPL> U;
#pragma unroll
    for (int k = 0; k < LPL; ++k) U.v[k] = (valid[k] ? an1.v[k] : kInfF) + X.v[k];
    const float u = wave_min(local_min(U));
    const Vec<LPL> m = minconv<LPL, DC, PE>(U, u, w, w2, L1v, L2v, a.delta, cc);
#pragma unroll
    for (int k = 0; k < LPL; ++k) X.v[k] = valid[k] ? m.v[k] - u : 0.f;
    u_lds[i] = u;  // every lane writes the same value
  };
  // Label (and, with GAP, the min-marginal gap) of node i from F_i + Bh_i.
  // Every lane stores the same value to the same address (one write, no
  // branch).
  auto decide = [&](int i, const Vec<LPL>& F, const Vec<LPL>& b) {
    const long p = pbase + (long)i * a.pns;
    float best = kInfF;
#pragma unroll
    for (int k = 0; k < LPL; ++k) best = fminf(best, valid[k] ? F.v[k] + b.v[k] : kInfF);
    const float gbest = wave_min(best);
    const float thr = gbest + kTieEps * fmaxf(1.f, fabsf(gbest));
    int bj = LPL;
#pragma unroll
    for (int k = LPL - 1; k >= 0; --k)
      if (valid[k] && F.v[k] + b.v[k] <= thr) bj = k;
    const unsigned long long ball = __ballot(bj < LPL);
    const int first = ball ? __builtin_ctzll(ball) : 0;  // empty only for NaN costs
    const int label = __builtin_amdgcn_readlane(lane * LPL + (bj < LPL ? bj : 0), first);
    if constexpr (GAP) {
      float second = kInfF;
#pragma unroll
      for (int k = 0; k < LPL; ++k) {
        const float mm = valid[k] ? F.v[k] + b.v[k] : kInfF;
        if (lane * LPL + k != label) second = fminf(second, mm);
      }
      a.gap[p] = wave_min(second) - gbest;
    }
    a.labels[p] = label;
  };

  // Phase 1: waves 0 and 1 walk to the midpoint, storing their messages.
  if (wv == 0) {
    X = ld_node<LPL, FULL>(Ab, lane, D, kInfF);  // F_0 = A_0
    if (mid > 0) st_node<LPL, FULL>(msg, lane, D, X);
    ring_pass<LPL, FULL, P, false, PE>(rA, strideB, rM, w_lds, w2_lds, n, D, lane, 1, mid - 1, 1, 0, -1,
                                       [&](int i, const Vec<LPL>& an, const Vec<LPL>&, float w, float w2) {
                                         fwd(an, w, w2);
                                         st_buf<LPL, FULL>(rM, i * D * 4, lane, D, X);
                                       });
  } else if (wv == 1) {
#pragma unroll
    for (int k = 0; k < LPL; ++k) X.v[k] = 0.f;  // Bh_{n-1} = 0
    st_node<LPL, FULL>(msg + (long)(n - 1) * D, lane, D, X);
    ring_pass<LPL, FULL, P, false, PE>(rA, strideB, rM, w_lds, w2_lds, n, D, lane, n - 2, n - 1 - mid, -1, 1, 0,
                                       [&](int i, const Vec<LPL>& an1, const Vec<LPL>&, float w, float w2) {
                                         bwd(i, an1, w, w2);
                                         st_buf<LPL, FULL>(rM, i * D * 4, lane, D, X);
                                       });
  }
  __threadfence_block();
  __syncthreads();  // the phase-1 messages are visible to the decision waves

  if (wv < 2) {
    // Phase 2, message waves.  Step j: wave 0 forms F at node mid + j (from
    // A_{mid+j} and w_{mid+j-1}; F_0 = A_0 is X already when mid = 0), wave 1
    // Bh at node mid-1-j (from A_{mid-j} and w_{mid-1-j}).  A runs P steps ahead
    // in a register ring; the block's weights and refill offsets sit one per
    // lane (v_readlane per step).
    const int len = wv == 0 ? len0 : len1;
    float* rbase = ring + wv * 2 * K * kNS;
    auto anode = [&](int j) { return wv == 0 ? min(mid + j, n - 1) : max(mid - j, 0); };
    auto wnode = [&](int j) { return max(min(wv == 0 ? mid + j - 1 : mid - 1 - j, n - 2), 0); };
    Vec<LPL> ra[P];
#pragma unroll
    for (int t = 0; t < P; ++t) ra[t] = ld_buf<LPL, FULL>(rA, anode(t) * strideB, lane, D, kInfF);
    for (int jb = 0; jb < S * K; jb += P) {
      const float wl = w_lds[wnode(jb + lane)];
      const float wl2 = PE ? w2_lds[wnode(jb + lane)] : 0.f;
      const int ao = anode(jb + P + lane) * strideB;
      const int r0 = jb / K;
#pragma unroll
      for (int t = 0; t < P; ++t) {
        const int j = jb + t;
        const Vec<LPL> an = fresh(ra[t]);
        __builtin_amdgcn_sched_barrier(kSchedValu);
        ra[t] = ld_buf<LPL, FULL>(rA, __builtin_amdgcn_readlane(ao, t), lane, D, kInfF);
        if (j < len) {
          if (wv == 0) {
            if (j > 0 || mid > 0) fwd(an, lane_weight(wl, t), PE ? lane_weight(wl2, t) : 0.f);
          } else {
            bwd(mid - 1 - j, an, lane_weight(wl, t), PE ? lane_weight(wl2, t) : 0.f);
          }
          st_ring<LPL>(rbase + (((r0 + t / K) & 1) * K + t % K) * kNS, lane, X);
        }
        if ((t + 1) % K == 0 && r0 + t / K < S) lds_barrier();  // end of round r0 + t / K
      }
    }
    lds_barrier();  // round S (the decision waves' last segment)
    if (wv == 1) {
      // Chain minimum: sum of the backward minima + min(A_0 + Bh_0).
      const Vec<LPL> a0 = ld_node<LPL, FULL>(Ab, lane, D, kInfF);
      Vec<LPL> M0;
#pragma unroll
      for (int k = 0; k < LPL; ++k) M0.v[k] = a0.v[k] + X.v[k];
      const double Lambda = chain_total(u_lds, n, lane) + (double)wave_min(local_min(M0));
      if (lane == 0) a.lam[c] = Lambda;
    }
  } else {
    // Phase 2, decision waves: wave 2 takes node mid + j with F from wave 0's
    // ring and Bh from the scratch volume, wave 3 node mid-1-j with Bh from
    // wave 1's ring and F from the scratch volume.  The scratch messages of the
    // next segment are loaded while this one is decided.
    const int h = wv - 2;
    const int len = h == 0 ? len0 : len1;
    const float* rbase = ring + h * 2 * K * kNS;
    auto node_of = [&](int j) { return h == 0 ? mid + j : mid - 1 - j; };
    auto load_other = [&](int s, Vec<LPL> (&o)[K]) {
#pragma unroll
      for (int t = 0; t < K; ++t) {
        const int i = min(max(node_of(s * K + t), 0), n - 1);
        o[t] = ld_buf<LPL, FULL>(rM, i * D * 4, lane, D, 0.f);
      }
    };
    auto decide_seg = [&](int s, const Vec<LPL> (&o)[K]) {
      const float* slot = rbase + (s & 1) * K * kNS;
#pragma unroll
      for (int t = 0; t < K; ++t) {
        const int j = s * K + t;
        if (j < len) {
          const Vec<LPL> mine = ld_ring<LPL>(slot + t * kNS, lane);
          if (h == 0)
            decide(node_of(j), mine, o[t]);
          else
            decide(node_of(j), o[t], mine);
        }
      }
    };
    Vec<LPL> OX[K], OY[K];
    load_other(0, OX);
    lds_barrier();  // round 0: segment 0 is written
    for (int r = 1; r <= S;) {
      load_other(min(r, S - 1), OY);
      decide_seg(r - 1, OX);
      lds_barrier();
      if (++r > S) break;
      load_other(min(r, S - 1), OX);
      decide_seg(r - 1, OY);
      lds_barrier();
      ++r;
    }
  }
}

// Compute units of the current device (cached per device).
static int cu_count() {
  static int cus[64] = {0};
  int dev = 0;
  (void)hipGetDevice(&dev);
  if (dev < 0 || dev >= 64) return 256;
  if (!cus[dev]) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || v <= 0) v = 256;
    cus[dev] = v;
  }
  return cus[dev];
}

constexpr size_t kLdsPerCu = 160 * 1024;

// LDS of one row chain with the message wave: b, w, u, [w2], v floats rounded
// to 16 bytes, then the message ring.
template <int LPL, bool PE>
static size_t row_seg_lds(int n) {
  return sizeof(float) * ((((PE ? 6 : 5) * (size_t)n + 3) & ~(size_t)3) +
                          (size_t)kRowRing * seg_len<LPL>() * 64 * LPL);
}

template <int LPL, int MODE, int P, bool FULL, int DC, bool SEG, bool NANCHK, bool PE>
static void launch_sweep(ChainArgs& a, size_t lds, hipStream_t st) {
  constexpr bool kRowSeg = MODE != kColsMinorant && SEG;
  int cpw = 1;
  if constexpr (kRowSeg) {
    lds = row_seg_lds<LPL, PE>(a.n);
    const int cus = cu_count();
    if (a.n_chains > cus && a.n_chains <= 2 * cus && 2 * lds <= kLdsPerCu) cpw = 2;  // one workgroup per CU
    lds *= cpw;
  }
  if (lds > 64 * 1024)  // long chains: allow the LDS beyond the default 64 KB
    (void)hipFuncSetAttribute((const void*)k_chain<LPL, MODE, P, FULL, DC, SEG, NANCHK, PE>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  k_chain<LPL, MODE, P, FULL, DC, SEG, NANCHK, PE><<<(a.n_chains + cpw - 1) / cpw, kRowSeg ? 128 * cpw : 64, lds, st>>>(a);
}

// Register-ring depth of the row sweeps with a message wave: three segments.
template <int LPL>
constexpr int row_seg_p() {
  return kRowRing * seg_len<LPL>();
}

template <int LPL, int MODE, int P, bool FULL, int DC, bool PE>
static void launch_one(ChainArgs& a, size_t lds, bool seg, hipStream_t st) {
  if constexpr (MODE == kRowsDecode) {
    // w (and w2), u, then the two halves' two-segment message rings at a 16-byte boundary
    const size_t lds2 = sizeof(float) * ((((PE ? 3 : 2) * (size_t)a.n + 3) & ~(size_t)3) +
                                         (size_t)2 * 2 * seg_len<LPL>() * 64 * LPL);
    constexpr int PD = row_seg_p<LPL>();
    if (lds2 > 64 * 1024) {
      (void)hipFuncSetAttribute((const void*)k_decode4<LPL, PD, FULL, DC, true, PE>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds2);
      (void)hipFuncSetAttribute((const void*)k_decode4<LPL, PD, FULL, DC, false, PE>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds2);
    }
    if (a.gap)
      k_decode4<LPL, PD, FULL, DC, true, PE><<<a.n_chains, 256, lds2, st>>>(a);
    else
      k_decode4<LPL, PD, FULL, DC, false, PE><<<a.n_chains, 256, lds2, st>>>(a);
  } else if constexpr (MODE == kColsMinorant) {
    if (seg)
      launch_sweep<LPL, MODE, P, FULL, DC, true, false, PE>(a, lds, st);
    else
      launch_sweep<LPL, MODE, P, FULL, DC, false, false, PE>(a, lds, st);
  } else if (seg) {
    if (a.nan_flags)
      launch_sweep<LPL, MODE, row_seg_p<LPL>(), FULL, DC, true, true, PE>(a, lds, st);
    else
      launch_sweep<LPL, MODE, row_seg_p<LPL>(), FULL, DC, true, false, PE>(a, lds, st);
  } else if (a.nan_flags) {
    launch_sweep<LPL, MODE, P, FULL, DC, false, true, PE>(a, lds, st);
  } else {
    launch_sweep<LPL, MODE, P, FULL, DC, false, false, PE>(a, lds, st);
  }
}

template <int MODE, int LPL, int P>
static void launch_lpl(ChainArgs& a, size_t lds, bool seg, hipStream_t st) {
  if (a.w2) {  // per-edge penalties
    if (a.D == 64 * LPL && a.delta == 2) launch_one<LPL, MODE, P, true, 2, true>(a, lds, seg, st);
    else launch_one<LPL, MODE, P, false, 0, true>(a, lds, seg, st);
  } else if (a.D == 64 * LPL && a.delta == 2) {
    launch_one<LPL, MODE, P, true, 2, false>(a, lds, seg, st);
  } else if (a.D == 64 * LPL) {
    launch_one<LPL, MODE, P, true, 0, false>(a, lds, seg, st);
  } else {
    launch_one<LPL, MODE, P, false, 0, false>(a, lds, seg, st);
  }
}

// Prefetch depth (nodes in flight per stream) by labels per lane; the row
// sweeps have only H chains, so they need a deeper ring than the columns to
// keep enough bytes in flight.
#ifndef CCS_CRF_P_ROWS
#define CCS_CRF_P_ROWS 32
#endif
#ifndef CCS_CRF_P_COLS
#define CCS_CRF_P_COLS 16
#endif

template <int MODE>
static int launch_mode(ChainArgs& a, bool seg, hipStream_t st) {
  const size_t lds = (sizeof(double) + (a.w2 ? 3 : 2) * sizeof(float)) * (size_t)a.n;
  constexpr int P2 = MODE == kColsMinorant ? CCS_CRF_P_COLS : CCS_CRF_P_ROWS;
  if (a.D <= 64) launch_lpl<MODE, 1, P2>(a, lds, seg, st);
  else if (a.D <= 128) launch_lpl<MODE, 2, P2>(a, lds, seg, st);
  else launch_lpl<MODE, 4, (P2 / 2 > 4 ? P2 / 2 : 4)>(a, lds, seg, st);
  CCS_LAUNCH_CHECK();
  return CCS_OK;
}

// Volumes from which the column sweeps recompute messages from checkpoints.
// Below it the sweep is latency-bound and the recomputation only adds steps;
// above it the sweep is HBM-bound and the recomputation removes two of the
// seven volume passes (measured: 0.304 -> 0.228 ms at 1242x375x128, 1.93 ->
// 1.55 ms at 1488x992x256; 0.056 -> 0.060 ms at 300x200x64).
constexpr long kSegMinBytes = 64l << 20;

// The threshold, overridable by the environment variable CCS_CRF_SEG_MIN_BYTES
// (read once per process; tests use it to run the checkpointed scheme on small
// volumes and compare it bit for bit with the store-everything one).
static long seg_min_bytes() {
  static const long v = [] {
    const char* e = getenv("CCS_CRF_SEG_MIN_BYTES");
    return e ? atol(e) : kSegMinBytes;
  }();
  return v;
}

int launch_chain(ChainMode mode, ChainArgs& a, hipStream_t st) {
  if (mode != kColsMinorant) {  // row chains are addressed with 32-bit byte offsets
    const long extent = ((long)(a.n - 1) * a.pns * a.D + a.D) * (long)sizeof(float);
    const long msg = (long)a.n * a.D * (long)sizeof(float);
    if (extent > 0x7fffffffL || msg > 0x7fffffffL)
      return fail(CCS_ERR_ARG, "row chain of %d nodes x %d labels exceeds 2 GiB", a.n, a.D);
  }
  bool seg = false;
  const long vol = (long)a.n_chains * a.n * a.D * (long)sizeof(float);
  if (mode == kColsMinorant && kSeg > 0) {
    // A segment's nodes are addressed with 32-bit byte offsets from its first node.
    const long span = ((long)(kSeg - 1) * a.pns * a.D + a.D) * (long)sizeof(float);
    seg = span <= 0x7fffffffL && vol >= seg_min_bytes();
  }
  // Row sweeps take the message wave at the same volumes (their chains are
  // addressed within the 2 GiB checked above).
  // The message wave also needs one chain's LDS to fit a CU.
  const bool pe = a.w2 != nullptr;
  const size_t row_lds = a.D <= 64    ? (pe ? row_seg_lds<1, true>(a.n) : row_seg_lds<1, false>(a.n))
                         : a.D <= 128 ? (pe ? row_seg_lds<2, true>(a.n) : row_seg_lds<2, false>(a.n))
                                      : (pe ? row_seg_lds<4, true>(a.n) : row_seg_lds<4, false>(a.n));
  const bool row_seg = kSeg > 0 && vol >= seg_min_bytes() && row_lds <= kLdsPerCu;
  switch (mode) {
    case kRowsMinorant: return launch_mode<kRowsMinorant>(a, row_seg, st);
    case kColsMinorant: return launch_mode<kColsMinorant>(a, seg, st);
    case kRowsDecode: return launch_mode<kRowsDecode>(a, false, st);
  }
  return fail(CCS_ERR_ARG, "unknown chain mode %d", (int)mode);
}

}  // namespace ccs
