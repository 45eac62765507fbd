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

template <int LPL, int MODE>
__global__ __launch_bounds__(64) void k_chain(ChainArgs a) {
  extern __shared__ double b_lds[];  // b_i = offset of the true backward message at node i
  const int c = blockIdx.x;
  const int lane = threadIdx.x;
  const int n = a.n, D = a.D;
  const long pbase = (long)c * a.pcs;
  float* __restrict__ msg = a.msg + (long)c * n * D;
  bool valid[LPL];
#pragma unroll
  for (int j = 0; j < LPL; ++j) valid[j] = lane * LPL + j < D;

  auto theta_at = [&](int i, Vec<LPL>& th, Vec<LPL>& av) {
    const long off = (pbase + (long)i * a.pns) * D;
    load_node<LPL>(a.A + off, lane, D, av, kInfF);
    if constexpr (MODE == kColsMinorant) {
      Vec<LPL> cv;
      load_node<LPL>(a.cost + off, lane, D, cv, kInfF);
#pragma unroll
      for (int j = 0; j < LPL; ++j) th.v[j] = valid[j] ? cv.v[j] - av.v[j] : kInfF;
    } else {
#pragma unroll
      for (int j = 0; j < LPL; ++j) th.v[j] = valid[j] ? av.v[j] : kInfF;
    }
  };
  auto weight_at = [&](int i) { return a.w[pbase + (long)i * a.pns]; };

  // ---- backward pass: Bh[n-1] = 0, Bh[i] = minconv(theta[i+1] + Bh[i+1]) - min(...)
  Vec<LPL> bh;
#pragma unroll
  for (int j = 0; j < LPL; ++j) bh.v[j] = 0.f;
  store_node<LPL>(msg + (long)(n - 1) * D, lane, D, bh);
  double bsum = 0.0;
  if (lane == 0) b_lds[n - 1] = 0.0;
  Vec<LPL> th_next, a_tmp;
  if (n >= 2) theta_at(n - 1, th_next, a_tmp);
  for (int i = n - 2; i >= 0; --i) {
    Vec<LPL> th_cur, a_cur;
    if (i >= 0) theta_at(i, th_cur, a_cur);  // prefetch theta[i] for the next step
    const float wi = weight_at(i);
    Vec<LPL> U;
#pragma unroll
    for (int j = 0; j < LPL; ++j) U.v[j] = th_next.v[j] + bh.v[j];
    const float u = wave_min(local_min(U));
    const Vec<LPL> m = minconv(U, u, wi, a.L1, a.L2, a.delta);
#pragma unroll
    for (int j = 0; j < LPL; ++j) bh.v[j] = valid[j] ? m.v[j] - u : 0.f;
    store_node<LPL>(msg + (long)i * D, lane, D, bh);
    bsum += (double)u;
    if (lane == 0) b_lds[i] = bsum;
    th_next = th_cur;
  }
  // Here th_next holds theta[0] (or theta[n-1] == theta[0] when n == 1), bh holds Bh[0].
  Vec<LPL> th0, a0;
  if (n >= 2) {
    th0 = th_next;
  } else {
    theta_at(0, th0, a0);
  }
  __syncthreads();
  const double b0 = b_lds[0];
  Vec<LPL> M0;
#pragma unroll
  for (int j = 0; j < LPL; ++j) M0.v[j] = th0.v[j] + bh.v[j];
  const double Lambda = b0 + (double)wave_min(local_min(M0));

  if constexpr (MODE == kRowsDecode) {
    // Forward min-sum F, label = argmin(F + Bh) (lowest label on ties).
    Vec<LPL> F = th0;
    for (int i = 0; i < n; ++i) {
      const long p = pbase + (long)i * a.pns;
      if (i > 0) {
        const float f = wave_min(local_min(F));
        const float wi = weight_at(i - 1);
        const Vec<LPL> m = minconv(F, f, wi, a.L1, a.L2, a.delta);
        Vec<LPL> th, av;
        theta_at(i, th, av);
#pragma unroll
        for (int j = 0; j < LPL; ++j) F.v[j] = valid[j] ? th.v[j] + m.v[j] - f : kInfF;
      }
      Vec<LPL> b;
      load_node<LPL>(msg + (long)i * D, lane, D, b, 0.f);
      float best = kInfF;
      int bj = 0;
#pragma unroll
      for (int j = 0; j < LPL; ++j) {
        const float mm = valid[j] ? F.v[j] + b.v[j] : kInfF;
        if (mm < best) {
          best = mm;
          bj = j;
        }
      }
      const float gbest = wave_min(best);
      const unsigned long long ball = __ballot(best == gbest);
      const int first = ball ? __builtin_ctzll(ball) : 0;  // empty only for NaN costs
      const int label = __builtin_amdgcn_readlane(lane * LPL + bj, first);
      if (a.gap) {
        float second = kInfF;
#pragma unroll
        for (int j = 0; j < LPL; ++j) {
          const float mm = valid[j] ? F.v[j] + b.v[j] : kInfF;
          if (lane * LPL + j != label) second = fminf(second, mm);
        }
        const float gsecond = wave_min(second);
        if (lane == 0) a.gap[p] = gsecond - gbest;
      }
      if (lane == 0) a.labels[p] = label;
    }
    if (lane == 0) a.lam[c] = Lambda;
    return;
  } else {
    // Forward pass: keep g_i = gamma (R + Bh) + excess, pass V = R - gamma (R + Bh) on.
    Vec<LPL> R = th0;
    double r = 0.0, kept = 0.0;
    Vec<LPL> a_cur;
    {
      Vec<LPL> tdummy;
      theta_at(0, tdummy, a_cur);
    }
    for (int i = 0; i + 1 < n; ++i) {
      const long off = (pbase + (long)i * a.pns) * D;
      Vec<LPL> b;
      load_node<LPL>(msg + (long)i * D, lane, D, b, 0.f);
      Vec<LPL> th_n, a_n;
      theta_at(i + 1, th_n, a_n);
      const float wi = weight_at(i);
      const double bi = b_lds[i];
      Vec<LPL> gt, V;
#pragma unroll
      for (int j = 0; j < LPL; ++j) {
        gt.v[j] = valid[j] ? kGamma * (R.v[j] + b.v[j]) : 0.f;
        V.v[j] = valid[j] ? R.v[j] - gt.v[j] : kInfF;
      }
      kept += kGammaD * (r + bi);
      const float v = wave_min(local_min(V));
      const Vec<LPL> m = minconv(V, v, wi, a.L1, a.L2, a.delta);
      Vec<LPL> mm;
#pragma unroll
      for (int j = 0; j < LPL; ++j) mm.v[j] = valid[j] ? m.v[j] : -kInfF;
      const Vec<LPL> nmax = near_max(mm, a.delta);
      const float wl1 = wi * a.L1;
      Vec<LPL> out;
#pragma unroll
      for (int j = 0; j < LPL; ++j) {
        const float e = valid[j] ? V.v[j] - fmaxf(m.v[j], nmax.v[j] - wl1) : 0.f;
        const float g = gt.v[j] + e;
        out.v[j] = MODE == kRowsMinorant ? a_cur.v[j] - g : a_cur.v[j] + g;
        R.v[j] = valid[j] ? th_n.v[j] + m.v[j] - v : kInfF;
      }
      store_node<LPL>(a.A + off, lane, D, out);
      r = (double)v + (1.0 - kGammaD) * r - kGammaD * bi;
      a_cur = a_n;
    }
    {
      const long off = (pbase + (long)(n - 1) * a.pns) * D;
      Vec<LPL> out;
#pragma unroll
      for (int j = 0; j < LPL; ++j)
        out.v[j] = MODE == kRowsMinorant ? a_cur.v[j] - R.v[j] : a_cur.v[j] + R.v[j];
      store_node<LPL>(a.A + off, lane, D, out);
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

template <int MODE>
static int launch_chain(ChainArgs& a, hipStream_t st) {
  const size_t lds = sizeof(double) * (size_t)a.n;
  if (a.D <= 64) {
    k_chain<1, MODE><<<a.n_chains, 64, lds, st>>>(a);
  } else if (a.D <= 128) {
    k_chain<2, MODE><<<a.n_chains, 64, lds, st>>>(a);
  } else {
    k_chain<4, MODE><<<a.n_chains, 64, lds, st>>>(a);
  }
  CCS_LAUNCH_CHECK();
  return CCS_OK;
}

}  // namespace ccs

using namespace ccs;

extern "C" int ccs_slackprop(ccs_handle_t h, const float* cost, const float* wx, const float* wy, int H, int W,
                             int D, float L1, float L2, int delta, int iters, int* labels, float* unaries_out,
                             double* sums, double* lower_bound) {
  CCS_CHECK(h && cost && wx && wy && labels, CCS_ERR_ARG, "ccs_slackprop: NULL argument");
  CCS_CHECK(H >= 1 && W >= 1 && D >= 2 && D <= 256, CCS_ERR_ARG,
            "ccs_slackprop: need H, W >= 1 and 2 <= D <= 256 (got H=%d W=%d D=%d)", H, W, D);
  CCS_CHECK(H <= 8192 && W <= 8192, CCS_ERR_UNSUPPORTED, "ccs_slackprop: chains longer than 8192");
  CCS_CHECK(iters >= 0 && iters <= 10000, CCS_ERR_ARG, "ccs_slackprop: bad iteration count %d", iters);
  CCS_CHECK(delta >= 1 && delta <= 64, CCS_ERR_ARG, "ccs_slackprop: delta must be in [1, 64] (got %d)", delta);
  CCS_CHECK(L1 >= 0.f && L2 >= L1, CCS_ERR_ARG,
            "ccs_slackprop: the min-sum form needs 0 <= L1 <= L2 (got L1=%g L2=%g)", L1, L2);
  const size_t vol = (size_t)H * W * D;
  int rc;
  if ((rc = h->crf_a.reserve(vol * sizeof(float)))) return rc;
  if ((rc = h->crf_msg.reserve(vol * sizeof(float)))) return rc;
  const int nmax = H > W ? H : W;
  const int nsums = 4 * iters + 2;
  if ((rc = h->crf_small.reserve(sizeof(double) * (2 * (size_t)nmax + nsums)))) return rc;
  float* A = (float*)h->crf_a.ptr;
  double* lam = (double*)h->crf_small.ptr;
  double* kept = lam + nmax;
  double* S = kept + nmax;            // S_0 .. S_{2 iters}
  double* K = S + 2 * iters + 1;      // K_0 .. K_{2 iters - 1}
  CCS_HIP(hipMemcpyAsync(A, cost, vol * sizeof(float), hipMemcpyDeviceToDevice, h->stream));

  ChainArgs rows{};
  rows.cost = cost; rows.A = A; rows.w = wx; rows.msg = (float*)h->crf_msg.ptr;
  rows.lam = lam; rows.kept = kept; rows.pcs = W; rows.pns = 1;
  rows.n_chains = H; rows.n = W; rows.D = D; rows.delta = delta; rows.L1 = L1; rows.L2 = L2;
  ChainArgs cols = rows;
  cols.w = wy; cols.pcs = 1; cols.pns = W; cols.n_chains = W; cols.n = H;

  for (int it = 0; it < iters; ++it) {
    if ((rc = launch_chain<kRowsMinorant>(rows, h->stream))) return rc;
    k_sum_f64<<<1, 256, 0, h->stream>>>(lam, H, S + 2 * it);
    k_sum_f64<<<1, 256, 0, h->stream>>>(kept, H, K + 2 * it);
    if ((rc = launch_chain<kColsMinorant>(cols, h->stream))) return rc;
    k_sum_f64<<<1, 256, 0, h->stream>>>(lam, W, S + 2 * it + 1);
    k_sum_f64<<<1, 256, 0, h->stream>>>(kept, W, K + 2 * it + 1);
  }
  ChainArgs dec = rows;
  dec.labels = labels;
  dec.gap = nullptr;
  if ((rc = launch_chain<kRowsDecode>(dec, h->stream))) return rc;
  k_sum_f64<<<1, 256, 0, h->stream>>>(lam, H, S + 2 * iters);
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
