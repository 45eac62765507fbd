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
// GPU mapping: one wavefront per chain, label s = lane * LPL + j held in VGPRs,
// DPP wave_shr/wave_shl for the label neighbours and a DPP tree for the minimum
// over labels.  The volume streams through HBM once per pass (node i is D
// contiguous floats), backward messages go to a scratch volume and the per-node
// float64 offsets to LDS.
#include <cmath>

#include "common.hpp"

namespace ccs {

constexpr float kGamma = 0.25f;
constexpr double kGammaD = 0.25;
constexpr float kInfF = __builtin_huge_valf();

enum ChainMode { kRowsMinorant = 0, kColsMinorant = 1, kRowsDecode = 2 };

struct ChainArgs {
  const float* cost;   // C (column mode)
  float* A;            // row unaries, updated in place
  const float* w;      // edge weight of the edge (node i, node i+1), indexed by pixel
  float* msg;          // backward messages, [chain][node][D]
  int* labels;         // decode output, indexed by pixel
  float* gap;          // decode: second best - best min-marginal (optional)
  double* lam;         // per chain: minimum chain energy
  double* kept;        // per chain: sum of kept constants
  long pcs, pns;       // pixel index = chain * pcs + node * pns
  int n_chains, n, D, delta;
  float L1, L2;
};

template <int LPL>
struct Vec {
  float v[LPL];
};

template <int LPL>
__device__ __forceinline__ void load_node(const float* __restrict__ base, int lane, int D, Vec<LPL>& out,
                                          float fill) {
  const int s0 = lane * LPL;
  if (D % LPL == 0 && s0 + LPL <= D) {  // 8/16-byte aligned: node offsets are multiples of D
    if constexpr (LPL == 4) {
      const float4 t = *reinterpret_cast<const float4*>(base + s0);
      out.v[0] = t.x; out.v[1] = t.y; out.v[2] = t.z; out.v[3] = t.w;
    } else if constexpr (LPL == 2) {
      const float2 t = *reinterpret_cast<const float2*>(base + s0);
      out.v[0] = t.x; out.v[1] = t.y;
    } else {
      out.v[0] = base[s0];
    }
  } else {
#pragma unroll
    for (int j = 0; j < LPL; ++j) out.v[j] = (s0 + j < D) ? base[s0 + j] : fill;
  }
}

template <int LPL>
__device__ __forceinline__ void store_node(float* __restrict__ base, int lane, int D, const Vec<LPL>& x) {
  const int s0 = lane * LPL;
  if (D % LPL == 0 && s0 + LPL <= D) {
    if constexpr (LPL == 4) {
      *reinterpret_cast<float4*>(base + s0) = make_float4(x.v[0], x.v[1], x.v[2], x.v[3]);
    } else if constexpr (LPL == 2) {
      *reinterpret_cast<float2*>(base + s0) = make_float2(x.v[0], x.v[1]);
    } else {
      base[s0] = x.v[0];
    }
  } else {
#pragma unroll
    for (int j = 0; j < LPL; ++j)
      if (s0 + j < D) base[s0 + j] = x.v[j];
  }
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

template <int LPL>
__device__ __forceinline__ float local_min(const Vec<LPL>& x) {
  float m = x.v[0];
#pragma unroll
  for (int j = 1; j < LPL; ++j) m = fminf(m, x.v[j]);
  return m;
}

// near(s) = min over 0 < |s - t| < delta of x(t)  (+inf if delta == 1).
template <int LPL>
__device__ __forceinline__ Vec<LPL> near_min(const Vec<LPL>& x, int delta) {
  if (delta == 2) {  // the common case, straight-line
    const Vec<LPL> l = shift_label(x, -1, kInfF), r = shift_label(x, +1, kInfF);
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

template <int LPL>
__device__ __forceinline__ Vec<LPL> near_max(const Vec<LPL>& x, int delta) {
  if (delta == 2) {
    const Vec<LPL> l = shift_label(x, -1, -kInfF), r = shift_label(x, +1, -kInfF);
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

// m(s) = min(U(s), near(U)(s) + w L1, min U + w L2).
template <int LPL>
__device__ __forceinline__ Vec<LPL> minconv(const Vec<LPL>& U, float umin, float w, float L1, float L2,
                                            int delta) {
  const Vec<LPL> nr = near_min(U, delta);
  const float far = umin + w * L2;
  const float wl1 = w * L1;
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
    if constexpr (LPL == 4) {
      *reinterpret_cast<float4*>(base + lane * 4) = make_float4(x.v[0], x.v[1], x.v[2], x.v[3]);
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

// Pipelined chain sweep.  Every input stream of a pass (unaries, cost, backward
// messages, edge weights) runs P nodes ahead of the node being processed: the
// main loop handles blocks of P nodes, unrolled, so ring slot j is a fixed set
// of registers and the load for node s + P is issued right after slot s is read
// (clamped to the last node, so it never leaves the chain).  The last partial
// block is an unrolled tail that issues no loads.
template <int LPL, int MODE, int P, bool FULL>
__global__ __launch_bounds__(64) void k_chain(ChainArgs a) {
  extern __shared__ double b_lds[];  // b_i = offset of the true backward message at node i
  constexpr bool kCols = MODE == kColsMinorant;
  const int c = blockIdx.x;
  const int lane = threadIdx.x;
  const int n = a.n, D = a.D;
  const long pbase = (long)c * a.pcs;
  const long nstride = a.pns * (long)D;  // floats between consecutive nodes
  const float* __restrict__ Ab = a.A + pbase * D;
  float* __restrict__ Aw = a.A + pbase * D;
  const float* __restrict__ Cb = kCols ? a.cost + pbase * D : nullptr;
  const float* __restrict__ wb = a.w + pbase;
  float* __restrict__ msg = a.msg + (long)c * n * D;
  bool valid[LPL];
#pragma unroll
  for (int j = 0; j < LPL; ++j) valid[j] = FULL || lane * LPL + j < D;

  auto theta_of = [&](const Vec<LPL>& av, const Vec<LPL>& cv) {
    Vec<LPL> th;
#pragma unroll
    for (int j = 0; j < LPL; ++j) th.v[j] = valid[j] ? (kCols ? cv.v[j] - av.v[j] : av.v[j]) : kInfF;
    return th;
  };

  // The chain's edge weights go to LDS once; each step reads its weight there.
  float* w_lds = reinterpret_cast<float*>(b_lds + n);
  for (int i = lane; i < n - 1; i += 64) w_lds[i] = wb[(long)i * a.pns];
  __syncthreads();

  Vec<LPL> ra[P], rc[P], rb[P];
  float rw[P];

  // ---- backward pass: step s handles i = n-2-s: Bh[i] = minconv(theta[i+1] + Bh[i+1]) - min(...)
  const int nb = n - 1;
  auto load_b = [&](int s, Vec<LPL>& av, Vec<LPL>& cv, float& w) {
    const int node = s < nb ? n - 1 - s : 1;  // clamped: beyond the chain re-read node 1
    av = ld_node<LPL, FULL>(Ab + node * nstride, lane, D, kInfF);
    if constexpr (kCols) cv = ld_node<LPL, FULL>(Cb + node * nstride, lane, D, kInfF);
    (void)w;
  };
  Vec<LPL> bh;
#pragma unroll
  for (int j = 0; j < LPL; ++j) bh.v[j] = 0.f;
  st_node<LPL, FULL>(msg + (long)(n - 1) * D, lane, D, bh);
  double bsum = 0.0;
  b_lds[n - 1] = 0.0;
  auto step_b = [&](int s, const Vec<LPL>& th, float) {
    const int i = n - 2 - s;
    const float wi = w_lds[i];
    Vec<LPL> U;
#pragma unroll
    for (int k = 0; k < LPL; ++k) U.v[k] = th.v[k] + bh.v[k];
    const float u = wave_min(local_min(U));
    const Vec<LPL> m = minconv(U, u, wi, a.L1, a.L2, a.delta);
#pragma unroll
    for (int k = 0; k < LPL; ++k) bh.v[k] = valid[k] ? m.v[k] - u : 0.f;
    st_node<LPL, FULL>(msg + (long)i * D, lane, D, bh);
    bsum += (double)u;
    b_lds[i] = bsum;  // every lane writes the same value
  };
  if (nb > 0) {
#pragma unroll
    for (int j = 0; j < P; ++j) load_b(j, ra[j], rc[j], rw[j]);
  }
  int s0 = 0;
  for (; s0 + P <= nb; s0 += P) {
#pragma unroll
    for (int j = 0; j < P; ++j) {
      const Vec<LPL> th = theta_of(ra[j], rc[j]);
      const float wi = rw[j];
      // Keep "read slot j" before "refill slot j": the refill then reuses the
      // slot's registers and the ring needs no copies at the loop back-edge.
      __builtin_amdgcn_sched_barrier(0);
      load_b(s0 + j + P, ra[j], rc[j], rw[j]);
      step_b(s0 + j, th, wi);
    }
  }
#pragma unroll
  for (int j = 0; j < P; ++j)
    if (s0 + j < nb) step_b(s0 + j, theta_of(ra[j], rc[j]), rw[j]);

  // bh holds Bh[0]; theta[0] is read directly.
  const Vec<LPL> a0 = ld_node<LPL, FULL>(Ab, lane, D, kInfF);
  Vec<LPL> c0;
  if constexpr (kCols) c0 = ld_node<LPL, FULL>(Cb, lane, D, kInfF);
  const Vec<LPL> th0 = theta_of(a0, c0);
  __syncthreads();
  Vec<LPL> M0;
#pragma unroll
  for (int k = 0; k < LPL; ++k) M0.v[k] = th0.v[k] + bh.v[k];
  const double Lambda = bsum + (double)wave_min(local_min(M0));

  if constexpr (MODE == kRowsDecode) {
    // Forward min-sum F, label = argmin(F + Bh) (lowest label on ties).
    auto load_d = [&](int i, Vec<LPL>& bv, Vec<LPL>& av, float& w) {
      const int node = i < n ? i : n - 1;
      bv = ld_node<LPL, FULL>(msg + (long)node * D, lane, D, 0.f);
      av = ld_node<LPL, FULL>(Ab + node * nstride, lane, D, kInfF);
      (void)w;
    };
    Vec<LPL> F = th0;
    auto step_d = [&](int i, const Vec<LPL>& b, const Vec<LPL>& an, float) {
      const long p = pbase + (long)i * a.pns;
      if (i > 0) {
        const float wv = w_lds[i - 1];
        const float f = wave_min(local_min(F));
        const Vec<LPL> m = minconv(F, f, wv, a.L1, a.L2, a.delta);
#pragma unroll
        for (int k = 0; k < LPL; ++k) F.v[k] = valid[k] ? an.v[k] + m.v[k] - f : kInfF;
      }
      float best = kInfF;
      int bj = 0;
#pragma unroll
      for (int k = 0; k < LPL; ++k) {
        const float mm = valid[k] ? F.v[k] + b.v[k] : kInfF;
        if (mm < best) {
          best = mm;
          bj = k;
        }
      }
      const float gbest = wave_min(best);
      const unsigned long long ball = __ballot(best == gbest);
      const int first = ball ? __builtin_ctzll(ball) : 0;  // empty only for NaN costs
      const int label = __builtin_amdgcn_readlane(lane * LPL + bj, first);
      if (a.gap) {
        float second = kInfF;
#pragma unroll
        for (int k = 0; k < LPL; ++k) {
          const float mm = valid[k] ? F.v[k] + b.v[k] : kInfF;
          if (lane * LPL + k != label) second = fminf(second, mm);
        }
        const float gsecond = wave_min(second);
        if (lane == 0) a.gap[p] = gsecond - gbest;
      }
      if (lane == 0) a.labels[p] = label;
    };
#pragma unroll
    for (int j = 0; j < P; ++j) load_d(j, rb[j], ra[j], rw[j]);
    int i0 = 0;
    for (; i0 + P <= n; i0 += P) {
#pragma unroll
      for (int j = 0; j < P; ++j) {
        const Vec<LPL> b = rb[j];
        const Vec<LPL> an = ra[j];
        const float wv = rw[j];
        __builtin_amdgcn_sched_barrier(0);
        load_d(i0 + j + P, rb[j], ra[j], rw[j]);
        step_d(i0 + j, b, an, wv);
      }
    }
#pragma unroll
    for (int j = 0; j < P; ++j)
      if (i0 + j < n) step_d(i0 + j, rb[j], ra[j], rw[j]);
    if (lane == 0) a.lam[c] = Lambda;
  } else {
    // Forward pass: keep g_i = gamma (R + Bh) + excess, pass V = R - gamma (R + Bh) on.
    auto load_f = [&](int i, Vec<LPL>& bv, Vec<LPL>& av, Vec<LPL>& cv, float& w) {
      const int node = i < nb ? i : nb - 1;
      bv = ld_node<LPL, FULL>(msg + (long)node * D, lane, D, 0.f);
      av = ld_node<LPL, FULL>(Ab + (node + 1) * nstride, lane, D, kInfF);
      if constexpr (kCols) cv = ld_node<LPL, FULL>(Cb + (node + 1) * nstride, lane, D, kInfF);
      (void)w;
    };
    Vec<LPL> R = th0;
    Vec<LPL> a_cur = a0;
    double r = 0.0, kept = 0.0;
    auto step_f = [&](int i, const Vec<LPL>& b, const Vec<LPL>& an, const Vec<LPL>& cn, float) {
      const double bi = b_lds[i];
      const float wi = w_lds[i];
      Vec<LPL> gt, V;
#pragma unroll
      for (int k = 0; k < LPL; ++k) {
        gt.v[k] = valid[k] ? kGamma * (R.v[k] + b.v[k]) : 0.f;
        V.v[k] = valid[k] ? R.v[k] - gt.v[k] : kInfF;
      }
      kept += kGammaD * (r + bi);
      const float v = wave_min(local_min(V));
      const Vec<LPL> m = minconv(V, v, wi, a.L1, a.L2, a.delta);
      Vec<LPL> mm;
#pragma unroll
      for (int k = 0; k < LPL; ++k) mm.v[k] = valid[k] ? m.v[k] : -kInfF;
      const Vec<LPL> nmax = near_max(mm, a.delta);
      const float wl1 = wi * a.L1;
      const Vec<LPL> thn = theta_of(an, cn);
      Vec<LPL> out;
#pragma unroll
      for (int k = 0; k < LPL; ++k) {
        const float e = valid[k] ? V.v[k] - fmaxf(m.v[k], nmax.v[k] - wl1) : 0.f;
        const float g = gt.v[k] + e;
        out.v[k] = MODE == kRowsMinorant ? a_cur.v[k] - g : a_cur.v[k] + g;
        R.v[k] = valid[k] ? thn.v[k] + m.v[k] - v : kInfF;
      }
      st_node<LPL, FULL>(Aw + i * nstride, lane, D, out);
      r = (double)v + (1.0 - kGammaD) * r - kGammaD * bi;
      a_cur = an;
    };
    if (nb > 0) {
#pragma unroll
      for (int j = 0; j < P; ++j) load_f(j, rb[j], ra[j], rc[j], rw[j]);
    }
    int i0 = 0;
    for (; i0 + P <= nb; i0 += P) {
#pragma unroll
      for (int j = 0; j < P; ++j) {
        const Vec<LPL> b = rb[j];
        const Vec<LPL> an = ra[j];
        const Vec<LPL> cn = rc[j];
        const float wi = rw[j];
        __builtin_amdgcn_sched_barrier(0);
        load_f(i0 + j + P, rb[j], ra[j], rc[j], rw[j]);
        step_f(i0 + j, b, an, cn, wi);
      }
    }
#pragma unroll
    for (int j = 0; j < P; ++j)
      if (i0 + j < nb) step_f(i0 + j, rb[j], ra[j], rc[j], rw[j]);
    {
      Vec<LPL> out;
#pragma unroll
      for (int k = 0; k < LPL; ++k)
        out.v[k] = MODE == kRowsMinorant ? a_cur.v[k] - R.v[k] : a_cur.v[k] + R.v[k];
      st_node<LPL, FULL>(Aw + (n - 1) * nstride, lane, D, out);
      kept += r;
    }
    if (lane == 0) {
      a.lam[c] = Lambda;
      a.kept[c] = kept;
    }
  }
}

// Deterministic sum of n doubles into out[0] (single block).
__global__ void k_sum_f64(const double* __restrict__ in, int n, double* __restrict__ out) {
  __shared__ double s[256];
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) acc += in[i];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = s[0];
}

// lb = sums[j] (+ kept[j-1] if j > 0).
__global__ void k_bound(const double* __restrict__ sums, int iters, double* __restrict__ lb) {
  const int j = 2 * iters;
  lb[0] = iters == 0 ? sums[0] : sums[j] + sums[2 * iters + 1 + (j - 1)];
}

__device__ __forceinline__ float rho(int d, float L1, float L2, int delta) {
  const int a = d < 0 ? -d : d;
  return a == 0 ? 0.f : (a < delta ? L1 : L2);
}

// Per-block float64 partial energy.
__global__ void k_energy_partial(const float* __restrict__ C, const float* __restrict__ wx,
                                 const float* __restrict__ wy, int H, int W, int D, const int* __restrict__ lab,
                                 float L1, float L2, int delta, double* __restrict__ part) {
  const long HW = (long)H * W;
  double acc = 0.0;
  for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < HW; p += (long)gridDim.x * blockDim.x) {
    const int x = (int)(p % W), y = (int)(p / W);
    const int l = lab[p];
    if (l < 0 || l >= D) {  // invalid label: the energy is undefined
      acc += __builtin_nan("");
      continue;
    }
    acc += (double)C[p * D + l];
    if (x + 1 < W) acc += (double)wx[p] * (double)rho(l - lab[p + 1], L1, L2, delta);
    if (y + 1 < H) acc += (double)wy[p] * (double)rho(l - lab[p + W], L1, L2, delta);
  }
  __shared__ double s[256];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

__global__ void k_ssvm_grad(const int* __restrict__ xhat, const int* __restrict__ xstar,
                            const float* __restrict__ wx, const float* __restrict__ wy, int H, int W, int D,
                            float L1, float L2, int delta, float* __restrict__ dwx, float* __restrict__ dwy,
                            double* __restrict__ part) {
  const long HW = (long)H * W;
  double d1 = 0.0, d2 = 0.0;
  for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < HW; p += (long)gridDim.x * blockDim.x) {
    const int x = (int)(p % W), y = (int)(p / W);
    const int ls = xstar[p], lh = xhat[p];
    float gx = 0.f, gy = 0.f;
    auto lev = [&](int a, int b, int which) {
      const int d = a > b ? a - b : b - a;
      return d == 0 ? 0.0 : (((which == 1) == (d < delta)) ? 1.0 : 0.0);
    };
    if (x + 1 < W) {
      d1 += (double)wx[p] * (lev(ls, xstar[p + 1], 1) - lev(lh, xhat[p + 1], 1));
      d2 += (double)wx[p] * (lev(ls, xstar[p + 1], 2) - lev(lh, xhat[p + 1], 2));
      gx = rho(ls - xstar[p + 1], L1, L2, delta) - rho(lh - xhat[p + 1], L1, L2, delta);
    }
    if (y + 1 < H) {
      d1 += (double)wy[p] * (lev(ls, xstar[p + W], 1) - lev(lh, xhat[p + W], 1));
      d2 += (double)wy[p] * (lev(ls, xstar[p + W], 2) - lev(lh, xhat[p + W], 2));
      gy = rho(ls - xstar[p + W], L1, L2, delta) - rho(lh - xhat[p + W], L1, L2, delta);
    }
    if (dwx) dwx[p] = gx;
    if (dwy) dwy[p] = gy;
  }
  __shared__ double s1[256], s2[256];
  s1[threadIdx.x] = d1;
  s2[threadIdx.x] = d2;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      s1[threadIdx.x] += s1[threadIdx.x + o];
      s2[threadIdx.x] += s2[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = s1[0];
    part[2 * blockIdx.x + 1] = s2[0];
  }
}

// dC[p][k] = [k == x*_p] - [k == xhat_p], written in full (coalesced).
__global__ void k_ssvm_unary(const int* __restrict__ xhat, const int* __restrict__ xstar, long HW, int D,
                             float* __restrict__ dC) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= HW * D) return;
  const long p = i / D;
  const int k = (int)(i % D);
  dC[i] = (k == xstar[p] ? 1.f : 0.f) - (k == xhat[p] ? 1.f : 0.f);
}

__global__ void k_sum_pairs(const double* __restrict__ part, int nb, double* __restrict__ out) {
  if (threadIdx.x < 2) {
    double s = 0.0;
    for (int b = 0; b < nb; ++b) s += part[2 * b + threadIdx.x];
    out[threadIdx.x] = s;
  }
}

template <int MODE, int LPL, int P>
static void launch_lpl(ChainArgs& a, size_t lds, hipStream_t st) {
  if (a.D == 64 * LPL)
    k_chain<LPL, MODE, P, true><<<a.n_chains, 64, lds, st>>>(a);
  else
    k_chain<LPL, MODE, P, false><<<a.n_chains, 64, lds, st>>>(a);
}

// Prefetch depth (nodes in flight per stream) by labels per lane; the row
// sweeps have only H chains, so they need a deeper ring than the columns to
// keep enough bytes in flight.
#ifndef CCS_CRF_P_ROWS
#define CCS_CRF_P_ROWS 16
#endif
#ifndef CCS_CRF_P_COLS
#define CCS_CRF_P_COLS 16
#endif

template <int MODE>
static int launch_chain(ChainArgs& a, hipStream_t st) {
  const size_t lds = (sizeof(double) + sizeof(float)) * (size_t)a.n;
  constexpr int P2 = MODE == kColsMinorant ? CCS_CRF_P_COLS : CCS_CRF_P_ROWS;
  if (a.D <= 64) launch_lpl<MODE, 1, 2 * P2>(a, lds, st);
  else if (a.D <= 128) launch_lpl<MODE, 2, P2>(a, lds, st);
  else launch_lpl<MODE, 4, (P2 / 2 > 4 ? P2 / 2 : 4)>(a, lds, st);
  CCS_LAUNCH_CHECK();
  return CCS_OK;
}

}  // namespace ccs

using namespace ccs;

static int check_crf_args(int H, int W, int D, float L1, float L2, int delta) {
  CCS_CHECK(H >= 1 && W >= 1 && D >= 2 && D <= 256, CCS_ERR_ARG,
            "slackprop: need H, W >= 1 and 2 <= D <= 256 (got H=%d W=%d D=%d)", H, W, D);
  CCS_CHECK(H <= 8192 && W <= 8192, CCS_ERR_UNSUPPORTED, "slackprop: chains longer than 8192");
  CCS_CHECK(delta >= 1 && delta <= 64, CCS_ERR_ARG, "slackprop: delta must be in [1, 64] (got %d)", delta);
  CCS_CHECK(L1 >= 0.f && L2 >= L1, CCS_ERR_ARG,
            "slackprop: the min-sum form needs 0 <= L1 <= L2 (got L1=%g L2=%g)", L1, L2);
  return CCS_OK;
}

extern "C" int ccs_slackprop_sweep(ccs_handle_t h, const float* cost, float* A, const float* wx, const float* wy,
                                   int H, int W, int D, float L1, float L2, int delta, int direction, int* labels,
                                   double* chain_sum, double* kept_sum) {
  CCS_CHECK(h && A, CCS_ERR_ARG, "ccs_slackprop_sweep: NULL argument");
  int rc = check_crf_args(H, W, D, L1, L2, delta);
  if (rc) return rc;
  CCS_CHECK(direction >= 0 && direction <= 2, CCS_ERR_ARG, "ccs_slackprop_sweep: direction must be 0, 1 or 2");
  CCS_CHECK(direction != 0 || wx, CCS_ERR_ARG, "ccs_slackprop_sweep: rows need wx");
  CCS_CHECK(direction != 1 || (wy && cost), CCS_ERR_ARG, "ccs_slackprop_sweep: columns need cost and wy");
  CCS_CHECK(direction != 2 || (wx && labels), CCS_ERR_ARG, "ccs_slackprop_sweep: decode needs wx and labels");
  const size_t vol = (size_t)H * W * D;
  if ((rc = h->crf_msg.reserve(vol * sizeof(float)))) return rc;
  const int nmax = H > W ? H : W;
  if ((rc = h->crf_small.reserve(sizeof(double) * 2 * (size_t)nmax))) return rc;
  double* lam = (double*)h->crf_small.ptr;
  double* kept = lam + nmax;
  ChainArgs a{};
  a.cost = cost; a.A = A; a.msg = (float*)h->crf_msg.ptr; a.lam = lam; a.kept = kept;
  a.D = D; a.delta = delta; a.L1 = L1; a.L2 = L2;
  if (direction == 1) {
    a.w = wy; a.pcs = 1; a.pns = W; a.n_chains = W; a.n = H;
  } else {
    a.w = wx; a.pcs = W; a.pns = 1; a.n_chains = H; a.n = W;
  }
  if (direction == 0) rc = launch_chain<kRowsMinorant>(a, h->stream);
  else if (direction == 1) rc = launch_chain<kColsMinorant>(a, h->stream);
  else {
    a.labels = labels;
    rc = launch_chain<kRowsDecode>(a, h->stream);
  }
  if (rc) return rc;
  if (chain_sum) k_sum_f64<<<1, 256, 0, h->stream>>>(lam, a.n_chains, chain_sum);
  if (kept_sum && direction != 2) k_sum_f64<<<1, 256, 0, h->stream>>>(kept, a.n_chains, kept_sum);
  CCS_LAUNCH_CHECK();
  return CCS_OK;
}

extern "C" int ccs_slackprop(ccs_handle_t h, const float* cost, const float* wx, const float* wy, int H, int W,
                             int D, float L1, float L2, int delta, int iters, int* labels, float* unaries_out,
                             double* sums, double* lower_bound) {
  CCS_CHECK(h && cost && wx && wy && labels, CCS_ERR_ARG, "ccs_slackprop: NULL argument");
  int rc = check_crf_args(H, W, D, L1, L2, delta);
  if (rc) return rc;
  CCS_CHECK(iters >= 0 && iters <= 10000, CCS_ERR_ARG, "ccs_slackprop: bad iteration count %d", iters);
  const size_t vol = (size_t)H * W * D;
  if ((rc = h->crf_a.reserve(vol * sizeof(float)))) return rc;
  if ((rc = h->crf_sums.reserve(sizeof(double) * (4 * (size_t)iters + 2)))) return rc;
  float* A = (float*)h->crf_a.ptr;
  double* S = (double*)h->crf_sums.ptr;  // S_0 .. S_{2 iters}
  double* K = S + 2 * iters + 1;         // K_0 .. K_{2 iters - 1}
  CCS_HIP(hipMemcpyAsync(A, cost, vol * sizeof(float), hipMemcpyDeviceToDevice, h->stream));
  for (int it = 0; it < iters; ++it) {
    if ((rc = ccs_slackprop_sweep(h, cost, A, wx, wy, H, W, D, L1, L2, delta, 0, nullptr, S + 2 * it, K + 2 * it)))
      return rc;
    if ((rc = ccs_slackprop_sweep(h, cost, A, wx, wy, H, W, D, L1, L2, delta, 1, nullptr, S + 2 * it + 1,
                                  K + 2 * it + 1)))
      return rc;
  }
  if ((rc = ccs_slackprop_sweep(h, cost, A, wx, wy, H, W, D, L1, L2, delta, 2, labels, S + 2 * iters, nullptr)))
    return rc;
  if (lower_bound) k_bound<<<1, 1, 0, h->stream>>>(S, iters, lower_bound);
  if (sums)
    CCS_HIP(hipMemcpyAsync(sums, S, sizeof(double) * (4 * (size_t)iters + 1), hipMemcpyDeviceToDevice, h->stream));
  if (unaries_out)
    CCS_HIP(hipMemcpyAsync(unaries_out, A, vol * sizeof(float), hipMemcpyDeviceToDevice, h->stream));
  CCS_LAUNCH_CHECK();
  return CCS_OK;
}

// Decode with the min-marginal gap (used by tests to classify label mismatches).
extern "C" int ccs_slackprop_decode_gap(ccs_handle_t h, const float* unaries, const float* wx, int H, int W, int D,
                                        float L1, float L2, int delta, int* labels, float* gap) {
  CCS_CHECK(h && unaries && wx && labels && gap, CCS_ERR_ARG, "ccs_slackprop_decode_gap: NULL argument");
  CCS_CHECK(H >= 1 && W >= 1 && D >= 2 && D <= 256, CCS_ERR_ARG, "ccs_slackprop_decode_gap: bad shape");
  const size_t vol = (size_t)H * W * D;
  int rc;
  if ((rc = h->crf_msg.reserve(vol * sizeof(float)))) return rc;
  if ((rc = h->crf_small.reserve(sizeof(double) * 2 * (size_t)H))) return rc;
  ChainArgs dec{};
  dec.A = const_cast<float*>(unaries); dec.w = wx; dec.msg = (float*)h->crf_msg.ptr;
  dec.labels = labels; dec.gap = gap; dec.lam = (double*)h->crf_small.ptr; dec.kept = dec.lam + H;
  dec.pcs = W; dec.pns = 1; dec.n_chains = H; dec.n = W; dec.D = D; dec.delta = delta; dec.L1 = L1; dec.L2 = L2;
  return launch_chain<kRowsDecode>(dec, h->stream);
}

extern "C" int ccs_slackprop_forward(ccs_handle_t h, const float* cost_volume, const float* wx, const float* wy,
                                     float* output, int in, int ic, int ih, int iw, float P1, float P2,
                                     int max_iter, float* lower_bound) {
  CCS_CHECK(in == 1, CCS_ERR_UNSUPPORTED, "ccs_slackprop_forward: only one sample (in == 1) is supported");
  CCS_CHECK(h && output, CCS_ERR_ARG, "ccs_slackprop_forward: NULL argument");
  const size_t hw = (size_t)ih * iw;
  int rc;
  if ((rc = h->misc.reserve(sizeof(int) * hw + sizeof(double)))) return rc;
  int* lab = (int*)h->misc.ptr;
  double* lb = (double*)((char*)h->misc.ptr + ((sizeof(int) * hw + 7) / 8) * 8);
  if ((rc = ccs_slackprop(h, cost_volume, wx, wy, ih, iw, ic, P1, P2, 2, max_iter, lab, nullptr, nullptr, lb)))
    return rc;
  // labels are written as float into the caller's output (stereo/operators/slackprop.cpp:133).
  if ((rc = ccs_remap(h, lab, ih, iw, 0.f, 1.f, output))) return rc;
  double lbh = 0.0;
  CCS_HIP(hipMemcpyAsync(&lbh, lb, sizeof(double), hipMemcpyDeviceToHost, h->stream));
  CCS_HIP(hipStreamSynchronize(h->stream));
  if (lower_bound) *lower_bound = (float)lbh;
  return CCS_OK;
}

extern "C" int ccs_crf_energy(ccs_handle_t h, const float* cost, const float* wx, const float* wy, int H, int W,
                              int D, const int* labels, float L1, float L2, int delta, double* energy) {
  CCS_CHECK(h && cost && wx && wy && labels && energy, CCS_ERR_ARG, "ccs_crf_energy: NULL argument");
  const int nb = 256;
  int rc;
  if ((rc = h->misc2.reserve(sizeof(double) * nb))) return rc;
  double* part = (double*)h->misc2.ptr;
  k_energy_partial<<<nb, 256, 0, h->stream>>>(cost, wx, wy, H, W, D, labels, L1, L2, delta, part);
  k_sum_f64<<<1, 256, 0, h->stream>>>(part, nb, energy);
  CCS_LAUNCH_CHECK();
  return CCS_OK;
}

extern "C" int ccs_crf_ssvm_grad(ccs_handle_t h, const int* xhat, const int* xstar, const float* wx,
                                 const float* wy, int H, int W, int D, float L1, float L2, int delta, float* dC,
                                 double* dL, float* dwx, float* dwy) {
  CCS_CHECK(h && xhat && xstar && wx && wy && dC && dL, CCS_ERR_ARG, "ccs_crf_ssvm_grad: NULL argument");
  const long HW = (long)H * W;
  const int nb = 256;
  int rc;
  if ((rc = h->misc2.reserve(sizeof(double) * 2 * nb))) return rc;
  double* part = (double*)h->misc2.ptr;
  k_ssvm_unary<<<ceil_div(HW * D, 256), 256, 0, h->stream>>>(xhat, xstar, HW, D, dC);
  k_ssvm_grad<<<nb, 256, 0, h->stream>>>(xhat, xstar, wx, wy, H, W, D, L1, L2, delta, dwx, dwy, part);
  k_sum_pairs<<<1, 64, 0, h->stream>>>(part, nb, dL);
  CCS_LAUNCH_CHECK();
  return CCS_OK;
}
