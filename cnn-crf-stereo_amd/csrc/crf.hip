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
#include <vector>

#include "crf_chain.hpp"

namespace ccs {


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

// Loss-augmented unaries of structured-SVM training: out = C - margin [l != t_p]
// (Hamming loss); pixels without ground truth (t outside [0, D)) keep C.
__global__ void k_loss_augment(const float* C, const int* __restrict__ t, long HW, int D, float margin, float* out) {
  const long n = HW * D;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const long p = i / D;
    const int l = (int)(i - p * D), tp = t[p];
    out[i] = C[i] - ((tp >= 0 && tp < D && l != tp) ? margin : 0.f);
  }
}

// Per-block float64 partial energy.  With a mask, pixels whose mask label is
// outside [0, D) (no ground truth) are left out: their unary and every edge
// with such an endpoint.
__global__ void k_energy_partial(const float* __restrict__ C, const float* __restrict__ wx,
                                 const float* __restrict__ wy, int H, int W, int D, const int* __restrict__ lab,
                                 const int* __restrict__ mask, float L1, float L2, int delta,
                                 double* __restrict__ part) {
  const long HW = (long)H * W;
  auto in = [&](long q) { return !mask || (mask[q] >= 0 && mask[q] < D); };
  double acc = 0.0;
  for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < HW; p += (long)gridDim.x * blockDim.x) {
    if (!in(p)) continue;
    const int x = (int)(p % W), y = (int)(p / W);
    const int l = lab[p];
    if (l < 0 || l >= D) {  // invalid label: the energy is undefined
      acc += __builtin_nan("");
      continue;
    }
    acc += (double)C[p * D + l];
    if (x + 1 < W && in(p + 1)) acc += (double)wx[p] * (double)rho(l - lab[p + 1], L1, L2, delta);
    if (y + 1 < H && in(p + W)) acc += (double)wy[p] * (double)rho(l - lab[p + W], L1, L2, delta);
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
  // Pixels whose ground truth x* is outside [0, D) are unlabelled: every edge
  // with such an endpoint is left out of both energies.
  auto lab = [&](long q) { return xstar[q] >= 0 && xstar[q] < D; };
  for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < HW; p += (long)gridDim.x * blockDim.x) {
    const int x = (int)(p % W), y = (int)(p / W);
    const int ls = xstar[p], lh = xhat[p];
    const bool lp = lab(p);
    float gx = 0.f, gy = 0.f;
    auto lev = [&](int a, int b, int which) {
      const int d = a > b ? a - b : b - a;
      return d == 0 ? 0.0 : (((which == 1) == (d < delta)) ? 1.0 : 0.0);
    };
    if (x + 1 < W && lp && lab(p + 1)) {
      d1 += (double)wx[p] * (lev(ls, xstar[p + 1], 1) - lev(lh, xhat[p + 1], 1));
      d2 += (double)wx[p] * (lev(ls, xstar[p + 1], 2) - lev(lh, xhat[p + 1], 2));
      gx = rho(ls - xstar[p + 1], L1, L2, delta) - rho(lh - xhat[p + 1], L1, L2, delta);
    }
    if (y + 1 < H && lp && lab(p + W)) {
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

// dC[p][k] = [k == x*_p] - [k == xhat_p], written in full (coalesced); 0 at
// unlabelled pixels (x*_p outside [0, D)).
__global__ void k_ssvm_unary(const int* __restrict__ xhat, const int* __restrict__ xstar, long HW, int D,
                             float* __restrict__ dC) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= HW * D) return;
  const long p = i / D;
  const int k = (int)(i % D), ts = xstar[p];
  dC[i] = (ts >= 0 && ts < D) ? (k == ts ? 1.f : 0.f) - (k == xhat[p] ? 1.f : 0.f) : 0.f;
}

__global__ void k_sum_pairs(const double* __restrict__ part, int nb, double* __restrict__ out) {
  if (threadIdx.x < 2) {
    double s = 0.0;
    for (int b = 0; b < nb; ++b) s += part[2 * b + threadIdx.x];
    out[threadIdx.x] = s;
  }
}

// Per-edge penalties: the pairwise volume pw (2 levels planes of H x W, the
// channels in the reference's order P1x, P1y, P2x, P2y, ...) re-laid out as the
// reference does before handing it to its solver (stereo/operators/
// slackprop.cpp:88-92): split the channels into (2, levels) with x/y fastest,
// flip the levels, and move x/y last: pen[l][y][x][c] = pw[c + 2 (levels-1-l)][y][x].
// Block b also writes bad[b] = 1 if one of its edges has P1 < 0 or P2 < P1
// (the min-sum form needs 0 <= P1 <= P2; P1 = pen[levels-1], P2 = pen[0]).
__global__ void k_pen_relayout(const float* __restrict__ pw, int levels, int H, int W, float* __restrict__ pen,
                               int* __restrict__ bad) {
  const long HW = (long)H * W;
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;  // pixel
  bool wrong = false;
  if (i < HW) {
    const int x = (int)(i % W), y = (int)(i / W);
    for (int l = 0; l < levels; ++l)
      for (int c = 0; c < 2; ++c) pen[((long)l * HW + i) * 2 + c] = pw[(long)(c + 2 * (levels - 1 - l)) * HW + i];
    for (int c = 0; c < 2; ++c) {
      const bool edge = c == 0 ? x + 1 < W : y + 1 < H;
      const float p1 = pw[(long)c * HW + i], p2 = pw[(long)(c + 2 * (levels - 1)) * HW + i];
      if (edge && !(p1 >= 0.f && p2 >= p1)) wrong = true;  // also catches NaN
    }
  }
  const int any = __syncthreads_or(wrong);
  if (threadIdx.x == 0) bad[blockIdx.x] = any;
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

// One half-step.  pen == nullptr: edge weights wx / wy with L1, L2; otherwise
// per-edge penalties pen (levels x H x W x 2: P1 = pen[levels-1], P2 = pen[0],
// x components for the row edges, y for the column edges) and wx, wy unused.
static int sweep(ccs_handle_t h, const float* cost, float* A, const float* wx, const float* wy, const float* pen,
                 int levels, int H, int W, int D, float L1, float L2, int delta, int direction, int* labels,
                 double* chain_sum, double* kept_sum) {
  CCS_CHECK(h && A, CCS_ERR_ARG, "ccs_slackprop_sweep: NULL argument");
  int rc = check_crf_args(H, W, D, pen ? 0.f : L1, pen ? 0.f : L2, delta);
  if (rc) return rc;
  CCS_CHECK(direction >= 0 && direction <= 3, CCS_ERR_ARG, "ccs_slackprop_sweep: direction must be 0, 1, 2 or 3");
  CCS_CHECK(pen || (direction != 0 && direction != 3) || wx, CCS_ERR_ARG, "ccs_slackprop_sweep: rows need wx");
  CCS_CHECK(direction == 2 || (cost && cost != A), CCS_ERR_ARG,
            "ccs_slackprop_sweep: row and column sweeps read cost, which must not be A");
  CCS_CHECK(direction != 1 || ((pen || wy) && cost), CCS_ERR_ARG, "ccs_slackprop_sweep: columns need cost and wy");
  CCS_CHECK(direction != 2 || ((pen || wx) && labels), CCS_ERR_ARG, "ccs_slackprop_sweep: decode needs wx and labels");
  const size_t vol = (size_t)H * W * D;
  if ((rc = h->crf_msg.reserve(vol * sizeof(float), h->stream))) return rc;
  const int nmax = H > W ? H : W;
  if ((rc = h->crf_small.reserve(sizeof(double) * 2 * (size_t)nmax, h->stream))) return rc;
  double* lam = (double*)h->crf_small.ptr;
  double* kept = lam + nmax;
  ChainArgs a{};
  if (direction == 3) {  // the first sweep of a run reads every cost: it flags NaN per row
    if ((rc = h->crf_nan.reserve(sizeof(int) * (size_t)H, h->stream))) return rc;
    a.nan_flags = (int*)h->crf_nan.ptr;
    h->nan_chains = H;
  }
  a.cost = cost; a.A = A; a.A_in = direction == 3 ? cost : A;
  a.msg = (float*)h->crf_msg.ptr; a.lam = lam; a.kept = kept;
  a.D = D; a.delta = delta; a.L1 = L1; a.L2 = L2;
  if (direction == 1) {
    a.w = wy; a.pcs = 1; a.pns = W; a.n_chains = W; a.n = H;
  } else {
    a.w = wx; a.pcs = W; a.pns = 1; a.n_chains = H; a.n = W;
  }
  if (pen) {
    const long plane = 2 * (long)H * W;
    const int comp = direction == 1 ? 1 : 0;  // y components couple the columns
    a.w = pen + (long)(levels - 1) * plane + comp;
    a.w2 = pen + comp;
    a.wmul = 2;
  }
  if (direction == 0 || direction == 3) rc = launch_chain(kRowsMinorant, a, h->stream);
  else if (direction == 1) rc = launch_chain(kColsMinorant, a, h->stream);
  else {
    a.labels = labels;
    rc = launch_chain(kRowsDecode, a, h->stream);
  }
  if (rc) return rc;
  if (chain_sum) k_sum_f64<<<1, 256, 0, h->stream>>>(lam, a.n_chains, chain_sum);
  if (kept_sum && direction != 2) k_sum_f64<<<1, 256, 0, h->stream>>>(kept, a.n_chains, kept_sum);
  CCS_LAUNCH_CHECK();
  return CCS_OK;
}

extern "C" int ccs_slackprop_sweep(ccs_handle_t h, const float* cost, float* A, const float* wx, const float* wy,
                                   int H, int W, int D, float L1, float L2, int delta, int direction, int* labels,
                                   double* chain_sum, double* kept_sum) {
  return sweep(h, cost, A, wx, wy, nullptr, 0, H, W, D, L1, L2, delta, direction, labels, chain_sum, kept_sum);
}

// The whole run (ccs_slackprop, or ccs_slackprop_pw with pen).
static int run(ccs_handle_t h, const float* cost, const float* wx, const float* wy, const float* pen, int levels,
               int H, int W, int D, float L1, float L2, int delta, int iters, int* labels, float* unaries_out,
               double* sums, double* lower_bound) {
  CCS_CHECK(h && cost && (pen || (wx && wy)) && labels, CCS_ERR_ARG, "ccs_slackprop: NULL argument");
  int rc = check_crf_args(H, W, D, pen ? 0.f : L1, pen ? 0.f : L2, delta);
  if (rc) return rc;
  CCS_CHECK(iters >= 0 && iters <= 10000, CCS_ERR_ARG, "ccs_slackprop: bad iteration count %d", iters);
  const size_t vol = (size_t)H * W * D;
  if ((rc = h->crf_a.reserve(vol * sizeof(float), h->stream))) return rc;
  if ((rc = h->crf_sums.reserve(sizeof(double) * (4 * (size_t)iters + 2), h->stream))) return rc;
  float* A = (float*)h->crf_a.ptr;
  double* S = (double*)h->crf_sums.ptr;  // S_0 .. S_{2 iters}
  double* K = S + 2 * iters + 1;         // K_0 .. K_{2 iters - 1}
  if (iters == 0) {
    CCS_HIP(hipMemcpyAsync(A, cost, vol * sizeof(float), hipMemcpyDeviceToDevice, h->stream));
    h->nan_chains = 0;  // no row sweep reads the costs
  }
  for (int it = 0; it < iters; ++it) {
    // The first row sweep reads the unaries from cost (A = cost is never copied).
    if ((rc = sweep(h, cost, A, wx, wy, pen, levels, H, W, D, L1, L2, delta, it == 0 ? 3 : 0, nullptr, S + 2 * it,
                    K + 2 * it)))
      return rc;
    if ((rc = sweep(h, cost, A, wx, wy, pen, levels, H, W, D, L1, L2, delta, 1, nullptr, S + 2 * it + 1,
                    K + 2 * it + 1)))
      return rc;
  }
  if ((rc = sweep(h, cost, A, wx, wy, pen, levels, H, W, D, L1, L2, delta, 2, labels, S + 2 * iters, nullptr)))
    return rc;
  if (lower_bound) k_bound<<<1, 1, 0, h->stream>>>(S, iters, lower_bound);
  if (sums)
    CCS_HIP(hipMemcpyAsync(sums, S, sizeof(double) * (4 * (size_t)iters + 1), hipMemcpyDeviceToDevice, h->stream));
  if (unaries_out)
    CCS_HIP(hipMemcpyAsync(unaries_out, A, vol * sizeof(float), hipMemcpyDeviceToDevice, h->stream));
  CCS_LAUNCH_CHECK();
  return CCS_OK;
}

extern "C" int ccs_slackprop(ccs_handle_t h, const float* cost, const float* wx, const float* wy, int H, int W,
                             int D, float L1, float L2, int delta, int iters, int* labels, float* unaries_out,
                             double* sums, double* lower_bound) {
  CCS_CHECK(h && wx && wy, CCS_ERR_ARG, "ccs_slackprop: NULL argument");
  return run(h, cost, wx, wy, nullptr, 0, H, W, D, L1, L2, delta, iters, labels, unaries_out, sums, lower_bound);
}

extern "C" int ccs_slackprop_pw(ccs_handle_t h, const float* cost, const float* pen, int levels, int H, int W, int D,
                                int delta, int iters, int* labels, float* unaries_out, double* sums,
                                double* lower_bound) {
  CCS_CHECK(h && pen, CCS_ERR_ARG, "ccs_slackprop_pw: NULL argument");
  CCS_CHECK(levels == 1 || levels == 2, CCS_ERR_UNSUPPORTED,
            "ccs_slackprop_pw: %d penalty levels (1: Potts, 2: P1 / P2 are implemented)", levels);
  return run(h, cost, nullptr, nullptr, pen, levels, H, W, D, 0.f, 0.f, delta, iters, labels, unaries_out, sums,
             lower_bound);
}

extern "C" int ccs_pairwise_penalties(ccs_handle_t h, const float* pw, int pw_channels, int H, int W, float* pen,
                                      int* bad) {
  CCS_CHECK(h && pw && pen && bad, CCS_ERR_ARG, "ccs_pairwise_penalties: NULL argument");
  CCS_CHECK(pw_channels == 2 || pw_channels == 4, CCS_ERR_UNSUPPORTED,
            "ccs_pairwise_penalties: %d channels (2: Potts, 4: P1 / P2 are implemented)", pw_channels);
  CCS_CHECK(H >= 1 && W >= 1, CCS_ERR_ARG, "ccs_pairwise_penalties: empty image");
  const long HW = (long)H * W;
  k_pen_relayout<<<ceil_div(HW, 256), 256, 0, h->stream>>>(pw, pw_channels / 2, H, W, pen, bad);
  CCS_LAUNCH_CHECK();
  return CCS_OK;
}

extern "C" int ccs_slackprop_forward_pw(ccs_handle_t h, const float* cost_volume, const float* pw_volume,
                                        float* output, int in, int ic, int ih, int iw, int pw_channels, int max_iter,
                                        float* lower_bound) {
  CCS_CHECK(in == 1, CCS_ERR_UNSUPPORTED, "ccs_slackprop_forward_pw: only one sample (in == 1) is supported");
  CCS_CHECK(h && cost_volume && pw_volume && output, CCS_ERR_ARG, "ccs_slackprop_forward_pw: NULL argument");
  CCS_CHECK(pw_channels == 2 || pw_channels == 4, CCS_ERR_UNSUPPORTED,
            "ccs_slackprop_forward_pw: %d pairwise channels (2: Potts, 4: P1 / P2 are implemented)", pw_channels);
  const size_t hw = (size_t)ih * iw;
  const int nbad = (int)ceil_div((long)hw, 256);
  // misc: labels | bound | per-block validation flags | penalties (levels x H x W x 2)
  const size_t off_lb = (sizeof(int) * hw + 7) / 8 * 8;
  const size_t off_bad = off_lb + sizeof(double);
  const size_t off_pen = (off_bad + sizeof(int) * nbad + 255) / 256 * 256;
  int rc;
  if ((rc = h->misc.reserve(off_pen + sizeof(float) * hw * pw_channels, h->stream))) return rc;
  char* base = (char*)h->misc.ptr;
  int* lab = (int*)base;
  double* lb = (double*)(base + off_lb);
  int* bad = (int*)(base + off_bad);
  float* pen = (float*)(base + off_pen);
  if ((rc = ccs_pairwise_penalties(h, pw_volume, pw_channels, ih, iw, pen, bad))) return rc;
  // delta = 2: the second level costs jumps of two labels or more (P1 for one label).
  if ((rc = ccs_slackprop_pw(h, cost_volume, pen, pw_channels / 2, ih, iw, ic, 2, max_iter, lab, nullptr, nullptr, lb)))
    return rc;
  // labels are written as float into the caller's output (stereo/operators/slackprop.cpp:103).
  if ((rc = ccs_remap(h, lab, ih, iw, 0.f, 1.f, output))) return rc;
  double lbh = 0.0;
  std::vector<int> badh(nbad);
  CCS_HIP(hipMemcpyAsync(&lbh, lb, sizeof(double), hipMemcpyDeviceToHost, h->stream));
  CCS_HIP(hipMemcpyAsync(badh.data(), bad, sizeof(int) * nbad, hipMemcpyDeviceToHost, h->stream));
  if ((rc = ccs_slackprop_status(h))) return rc;  // synchronises
  for (int b = 0; b < nbad; ++b)
    CCS_CHECK(!badh[b], CCS_ERR_ARG,
              "ccs_slackprop_forward_pw: the pairwise volume needs 0 <= P1 <= P2 on every edge (near pixel %d)",
              b * 256);
  if (lower_bound) *lower_bound = (float)lbh;
  return CCS_OK;
}

// Decode with the min-marginal gap (used by tests to classify label mismatches).
extern "C" int ccs_slackprop_decode_gap(ccs_handle_t h, const float* unaries, const float* wx, int H, int W, int D,
                                        float L1, float L2, int delta, int* labels, float* gap) {
  CCS_CHECK(h && unaries && wx && labels && gap, CCS_ERR_ARG, "ccs_slackprop_decode_gap: NULL argument");
  int rc = check_crf_args(H, W, D, L1, L2, delta);
  if (rc) return rc;
  const size_t vol = (size_t)H * W * D;
  if ((rc = h->crf_msg.reserve(vol * sizeof(float), h->stream))) return rc;
  if ((rc = h->crf_small.reserve(sizeof(double) * 2 * (size_t)H, h->stream))) return rc;
  ChainArgs dec{};
  dec.A = const_cast<float*>(unaries); dec.A_in = unaries; dec.w = wx; dec.msg = (float*)h->crf_msg.ptr;
  dec.labels = labels; dec.gap = gap; dec.lam = (double*)h->crf_small.ptr; dec.kept = dec.lam + H;
  dec.pcs = W; dec.pns = 1; dec.n_chains = H; dec.n = W; dec.D = D; dec.delta = delta; dec.L1 = L1; dec.L2 = L2;
  return launch_chain(kRowsDecode, dec, h->stream);
}

extern "C" int ccs_slackprop_status(ccs_handle_t h) {
  CCS_CHECK(h != nullptr, CCS_ERR_ARG, "ccs_slackprop_status: NULL handle");
  CCS_HIP(hipStreamSynchronize(h->stream));
  if (h->nan_chains <= 0) return CCS_OK;
  std::vector<int> f((size_t)h->nan_chains);
  CCS_HIP(hipMemcpy(f.data(), h->crf_nan.ptr, sizeof(int) * f.size(), hipMemcpyDeviceToHost));
  int rows = 0, first = -1;
  for (size_t i = 0; i < f.size(); ++i)
    if (f[i]) {
      if (first < 0) first = (int)i;
      ++rows;
    }
  CCS_CHECK(rows == 0, CCS_ERR_NAN, "slackprop: the cost volume holds NaN in %d row(s), the first at y = %d", rows,
            first);
  return CCS_OK;
}

extern "C" int ccs_slackprop_forward(ccs_handle_t h, const float* cost_volume, const float* wx, const float* wy,
                                     float* output, int in, int ic, int ih, int iw, float P1, float P2,
                                     int max_iter, float* lower_bound) {
  CCS_CHECK(in == 1, CCS_ERR_UNSUPPORTED, "ccs_slackprop_forward: only one sample (in == 1) is supported");
  CCS_CHECK(h && output, CCS_ERR_ARG, "ccs_slackprop_forward: NULL argument");
  const size_t hw = (size_t)ih * iw;
  int rc;
  if ((rc = h->misc.reserve(sizeof(int) * hw + sizeof(double), h->stream))) return rc;
  int* lab = (int*)h->misc.ptr;
  double* lb = (double*)((char*)h->misc.ptr + ((sizeof(int) * hw + 7) / 8) * 8);
  if ((rc = ccs_slackprop(h, cost_volume, wx, wy, ih, iw, ic, P1, P2, 2, max_iter, lab, nullptr, nullptr, lb)))
    return rc;
  // labels are written as float into the caller's output (stereo/operators/slackprop.cpp:133).
  if ((rc = ccs_remap(h, lab, ih, iw, 0.f, 1.f, output))) return rc;
  double lbh = 0.0;
  CCS_HIP(hipMemcpyAsync(&lbh, lb, sizeof(double), hipMemcpyDeviceToHost, h->stream));
  if ((rc = ccs_slackprop_status(h))) return rc;  // synchronises
  if (lower_bound) *lower_bound = (float)lbh;
  return CCS_OK;
}

extern "C" int ccs_slackprop_forward_volume_out(ccs_handle_t h, const float* cost_volume, const float* wx,
                                                const float* wy, float* output, int in, int ic, int ih, int iw,
                                                float P1, float P2, int max_iter, int delta, float* lower_bound) {
  CCS_CHECK(in == 1, CCS_ERR_UNSUPPORTED, "ccs_slackprop_forward_volume_out: only one sample (in == 1) is supported");
  CCS_CHECK(h && output, CCS_ERR_ARG, "ccs_slackprop_forward_volume_out: NULL argument");
  const size_t hw = (size_t)ih * iw;
  int rc;
  if ((rc = h->misc.reserve(sizeof(int) * hw + 2 * sizeof(double), h->stream))) return rc;
  int* lab = (int*)h->misc.ptr;
  double* lb = (double*)((char*)h->misc.ptr + ((sizeof(int) * hw + 7) / 8) * 8);
  // output receives the reparametrised volume, the solver's get_C1 (stereo/operators/slackprop.cpp:168).
  if ((rc = ccs_slackprop(h, cost_volume, wx, wy, ih, iw, ic, P1, P2, delta, max_iter, lab, output, nullptr, lb)))
    return rc;
  double lbh = 0.0;
  CCS_HIP(hipMemcpyAsync(&lbh, lb, sizeof(double), hipMemcpyDeviceToHost, h->stream));
  if ((rc = ccs_slackprop_status(h))) return rc;  // synchronises
  if (lower_bound) *lower_bound = (float)lbh;
  return CCS_OK;
}

extern "C" int ccs_crf_loss_augment(ccs_handle_t h, const float* cost, const int* target, long HW, int D,
                                    float margin, float* out) {
  CCS_CHECK(h && cost && target && out && HW >= 0 && D >= 1, CCS_ERR_ARG, "ccs_crf_loss_augment: bad argument");
  if (HW == 0) return CCS_OK;
  const long n = HW * D;
  const long nb = (n + 255) / 256;
  k_loss_augment<<<(unsigned)(nb < 65536 ? nb : 65536), 256, 0, h->stream>>>(cost, target, HW, D, margin, out);
  CCS_LAUNCH_CHECK();
  return CCS_OK;
}

static int crf_energy(ccs_handle_t h, const float* cost, const float* wx, const float* wy, int H, int W, int D,
                      const int* labels, const int* mask, float L1, float L2, int delta, double* energy) {
  const int nb = kReduceBlocks;
  int rc;
  if ((rc = h->misc2.reserve(sizeof(double) * nb, h->stream))) return rc;
  double* part = (double*)h->misc2.ptr;
  k_energy_partial<<<nb, 256, 0, h->stream>>>(cost, wx, wy, H, W, D, labels, mask, L1, L2, delta, part);
  k_sum_f64<<<1, 256, 0, h->stream>>>(part, nb, energy);
  CCS_LAUNCH_CHECK();
  return CCS_OK;
}

extern "C" int ccs_crf_energy(ccs_handle_t h, const float* cost, const float* wx, const float* wy, int H, int W,
                              int D, const int* labels, float L1, float L2, int delta, double* energy) {
  CCS_CHECK(h && cost && wx && wy && labels && energy, CCS_ERR_ARG, "ccs_crf_energy: NULL argument");
  return crf_energy(h, cost, wx, wy, H, W, D, labels, nullptr, L1, L2, delta, energy);
}

extern "C" int ccs_crf_energy_masked(ccs_handle_t h, const float* cost, const float* wx, const float* wy, int H,
                                     int W, int D, const int* labels, const int* mask, float L1, float L2, int delta,
                                     double* energy) {
  CCS_CHECK(h && cost && wx && wy && labels && mask && energy, CCS_ERR_ARG, "ccs_crf_energy_masked: NULL argument");
  return crf_energy(h, cost, wx, wy, H, W, D, labels, mask, L1, L2, delta, energy);
}

extern "C" int ccs_crf_ssvm_grad(ccs_handle_t h, const int* xhat, const int* xstar, const float* wx,
                                 const float* wy, int H, int W, int D, float L1, float L2, int delta, float* dC,
                                 double* dL, float* dwx, float* dwy) {
  CCS_CHECK(h && xhat && xstar && wx && wy && dC && dL, CCS_ERR_ARG, "ccs_crf_ssvm_grad: NULL argument");
  const long HW = (long)H * W;
  const int nb = kReduceBlocks;
  int rc;
  if ((rc = h->misc2.reserve(sizeof(double) * 2 * nb, h->stream))) return rc;
  double* part = (double*)h->misc2.ptr;
  k_ssvm_unary<<<ceil_div(HW * D, 256), 256, 0, h->stream>>>(xhat, xstar, HW, D, dC);
  k_ssvm_grad<<<nb, 256, 0, h->stream>>>(xhat, xstar, wx, wy, H, W, D, L1, L2, delta, dwx, dwy, part);
  k_sum_pairs<<<1, 64, 0, h->stream>>>(part, nb, dL);
  CCS_LAUNCH_CHECK();
  return CCS_OK;
}
