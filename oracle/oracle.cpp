/*
 * FP64 CPU oracle of the CNN-CRF stereo hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * This file is the checker: tests, the smoke check and the benchmark's CPU
 * baseline load it; the HIP product never does.  Each function restates one
 * reference operator in double precision (citations next to each function,
 * relative to the reference repository root).
 *
 * SlackProp: the reference calls slack_prop_2D_alg from an un-vendored library
 * (stereo/stereo.cpp:140-154, stereo/operators/slackprop.cpp:96-105).  Its
 * source is absent, so the CRF here is a restatement of the published dual
 * minorize-maximize method over row and column chains; the GPU implements this
 * same restatement.  Parity with the upstream library is unpinned; the
 * restatement itself is pinned by exact chain minima, invariants and brute force
 * in the tests.
 */
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <vector>

namespace {
constexpr double kInf = std::numeric_limits<double>::infinity();
// Relative tolerance under which decoded min-marginals count as tied (the GPU
// decode uses the same rule in float32).
constexpr double kTieEps = 1e-5;
}

extern "C" int or_reflect(int i, int n) {
  if (n == 1) return 0;
  const int period = 2 * n;
  i %= period;
  if (i < 0) i += period;
  return i < n ? i : period - 1 - i;
}

// stereo/stereo.cu:865-889: the padded pixel (x, y) samples the source at the
// normalised coordinate (x - p + 0.5) / W under mirror addressing, i.e. the
// source index reflected with the edge pixel repeated.
extern "C" void or_pad_symmetric(const double* in, int C, int H, int W, int p, double* out) {
  const int Ho = H + 2 * p, Wo = W + 2 * p;
  for (int c = 0; c < C; ++c)
    for (int y = 0; y < Ho; ++y) {
      const int sy = or_reflect(y - p, H);
      for (int x = 0; x < Wo; ++x)
        out[((long)c * Ho + y) * Wo + x] = in[((long)c * H + sy) * W + or_reflect(x - p, W)];
    }
}

// stereo/utils.cu:169-237 (makeRgbZeroMean + makeRgbUnitStd(isZeroMean=true)):
// subtract the channel mean, then divide by sqrt(sum x^2 / N) (population std).
// A constant channel (std == 0) is left unscaled instead of dividing by zero.
extern "C" void or_normalize(double* x, int C, long HW, double* mean_out, double* std_out) {
  for (int c = 0; c < C; ++c) {
    double* p = x + (long)c * HW;
    double s = 0.0;
    for (long i = 0; i < HW; ++i) s += p[i];
    const double mean = s / (double)HW;
    double q = 0.0;
    for (long i = 0; i < HW; ++i) {
      p[i] -= mean;
      q += p[i] * p[i];
    }
    const double sd = std::sqrt(q / (double)HW);
    const double inv = sd > 0.0 ? 1.0 / sd : 1.0;
    for (long i = 0; i < HW; ++i) p[i] *= inv;
    if (mean_out) mean_out[c] = mean;
    if (std_out) std_out[c] = sd;
  }
}

// stereo/stereo.cu:853-863.
extern "C" void or_rgb2gray(const double* rgb, int H, int W, double* gray) {
  const long n = (long)H * W;
  for (long i = 0; i < n; ++i)
    gray[i] = 0.2989 * rgb[i] + 0.5870 * rgb[n + i] + 0.1140 * rgb[2 * n + i];
}

// stereo/operators/convolution.cpp:108,149 (CUDNN_CONVOLUTION: the filter is
// flipped), pad 0, stride 1; bias stereo/operators/bias.cpp:58-66; tanh
// stereo/operators/activation.cpp:33-40.
extern "C" void or_conv2d(const double* x, int C, int H, int W, const double* w, const double* b,
                          int K, int KH, int KW, int act_tanh, double* y) {
  const int Ho = H - KH + 1, Wo = W - KW + 1;
#pragma omp parallel for schedule(dynamic)
  for (int k = 0; k < K; ++k) {
    double* yk = y + (long)k * Ho * Wo;
    const double bias = b ? b[k] : 0.0;
    for (long i = 0; i < (long)Ho * Wo; ++i) yk[i] = bias;
    for (int c = 0; c < C; ++c)
      for (int i = 0; i < KH; ++i)
        for (int j = 0; j < KW; ++j) {
          const double wt = w[(((long)k * C + c) * KH + (KH - 1 - i)) * KW + (KW - 1 - j)];
          for (int h = 0; h < Ho; ++h) {
            const double* xr = x + ((long)c * H + h + i) * W + j;
            double* yr = yk + (long)h * Wo;
            for (int ww = 0; ww < Wo; ++ww) yr[ww] += wt * xr[ww];
          }
        }
    if (act_tanh)
      for (long i = 0; i < (long)Ho * Wo; ++i) yk[i] = std::tanh(yk[i]);
  }
}

extern "C" void or_conv2d_bwd(const double* x, int C, int H, int W, const double* w, const double* b,
                              int K, int KH, int KW, int act_tanh, const double* dy,
                              double* dx, double* dw, double* db) {
  const int Ho = H - KH + 1, Wo = W - KW + 1;
  const long n_out = (long)K * Ho * Wo;
  std::vector<double> dz(n_out);
  if (act_tanh) {
    std::vector<double> a(n_out);
    or_conv2d(x, C, H, W, w, b, K, KH, KW, 1, a.data());
    for (long i = 0; i < n_out; ++i) dz[i] = dy[i] * (1.0 - a[i] * a[i]);
  } else {
    std::memcpy(dz.data(), dy, sizeof(double) * n_out);
  }
  if (db)
    for (int k = 0; k < K; ++k) {
      double s = 0.0;
      for (long i = 0; i < (long)Ho * Wo; ++i) s += dz[(long)k * Ho * Wo + i];
      db[k] = s;
    }
  if (dw) {
#pragma omp parallel for schedule(dynamic)
    for (int k = 0; k < K; ++k)
      for (int c = 0; c < C; ++c)
        for (int i = 0; i < KH; ++i)
          for (int j = 0; j < KW; ++j) {
            double s = 0.0;
            for (int h = 0; h < Ho; ++h) {
              const double* xr = x + ((long)c * H + h + i) * W + j;
              const double* zr = dz.data() + ((long)k * Ho + h) * Wo;
              for (int ww = 0; ww < Wo; ++ww) s += zr[ww] * xr[ww];
            }
            dw[(((long)k * C + c) * KH + (KH - 1 - i)) * KW + (KW - 1 - j)] = s;
          }
  }
  if (dx) {
#pragma omp parallel for schedule(dynamic)
    for (int c = 0; c < C; ++c) {
      double* dxc = dx + (long)c * H * W;
      for (long i = 0; i < (long)H * W; ++i) dxc[i] = 0.0;
      for (int k = 0; k < K; ++k)
        for (int i = 0; i < KH; ++i)
          for (int j = 0; j < KW; ++j) {
            const double wt = w[(((long)k * C + c) * KH + (KH - 1 - i)) * KW + (KW - 1 - j)];
            for (int h = 0; h < Ho; ++h) {
              double* xr = dxc + (long)(h + i) * W + j;
              const double* zr = dz.data() + ((long)k * Ho + h) * Wo;
              for (int ww = 0; ww < Wo; ++ww) xr[ww] += wt * zr[ww];
            }
          }
    }
  }
}

// stereo/operators/stereocorrelation.cu:111-142 (kCorrelation_NHWC); the
// disparity list is built at stereo/operators/stereocorrelation.cpp:83-85.
// Offsets falling outside the image on either side give 0 (the reference
// only tests the lower side; the upper side is defined here as 0 too).
extern "C" void or_correlation(const double* f0, const double* f1, int C, int H, int W,
                               const int* disp, int D, int rect_corr, double* out) {
  const long HW = (long)H * W;
#pragma omp parallel for schedule(static)
  for (int h = 0; h < H; ++h)
    for (int w = 0; w < W; ++w)
      for (int k = 0; k < D; ++k) {
        const int w2 = w + disp[k];
        double best = -100000.0;
        for (int yy = -rect_corr; yy <= rect_corr; ++yy) {
          const int h2 = h + yy;
          if (h2 < 0 || h2 >= H) continue;
          double r = 0.0;
          if (w2 >= 0 && w2 < W)
            for (int c = 0; c < C; ++c) r += f0[c * HW + (long)h * W + w] * f1[c * HW + (long)h2 * W + w2];
          if (r > best) best = r;
        }
        out[((long)h * W + w) * D + k] = best;
      }
}

// stereo/operators/stereocorrelation.cu:243-278 (kCorrelationGrad_NHWC).
extern "C" void or_correlation_bwd(const double* f0, const double* f1, int C, int H, int W,
                                   const int* disp, int D, const double* G, double* df0, double* df1) {
  const long HW = (long)H * W;
#pragma omp parallel for schedule(static)
  for (int c = 0; c < C; ++c)
    for (int h = 0; h < H; ++h)
      for (int w = 0; w < W; ++w) {
        double g0 = 0.0, g1 = 0.0;
        for (int k = 0; k < D; ++k) {
          const int wp = w + disp[k];
          if (wp >= 0 && wp < W) g0 += f1[c * HW + (long)h * W + wp] * G[((long)h * W + w) * D + k];
          const int wm = w - disp[k];
          if (wm >= 0 && wm < W) g1 += f0[c * HW + (long)h * W + wm] * G[((long)h * W + wm) * D + k];
        }
        df0[c * HW + (long)h * W + w] = g0;
        df1[c * HW + (long)h * W + w] = g1;
      }
}

// stereo/operators/softmax.cpp:30-38 (ACCURATE softmax over the channel = D
// axis of an NHWC tensor) followed by stereo/utils.cu:289-292 (negate).
extern "C" void or_neg_softmax(double* vol, long HW, int D) {
#pragma omp parallel for schedule(static)
  for (long p = 0; p < HW; ++p) {
    double* v = vol + p * D;
    double m = -kInf;
    for (int k = 0; k < D; ++k) m = std::max(m, v[k]);
    double s = 0.0;
    for (int k = 0; k < D; ++k) {
      v[k] = std::exp(v[k] - m);
      s += v[k];
    }
    for (int k = 0; k < D; ++k) v[k] = -v[k] / s;
  }
}

extern "C" double or_softmax_xent(const double* S, long HW, int D, const int* target, double* dS) {
  double loss = 0.0;
  for (long p = 0; p < HW; ++p) {
    const double* v = S + p * D;
    double m = -kInf;
    for (int k = 0; k < D; ++k) m = std::max(m, v[k]);
    double s = 0.0;
    for (int k = 0; k < D; ++k) s += std::exp(v[k] - m);
    const double lse = m + std::log(s);
    const int t = target[p];
    const bool valid = t >= 0 && t < D;  // other targets mark pixels without ground truth
    if (valid) loss += lse - v[t];
    if (dS)
      for (int k = 0; k < D; ++k)
        dS[p * D + k] = valid ? std::exp(v[k] - lse) - (k == t ? 1.0 : 0.0) : 0.0;
  }
  return loss;
}

// Separable Gaussian standing in for iu::filterGauss(sigma) at
// stereo/stereo.cu:1092 (library un-vendored): radius ceil(3 sigma),
// normalised taps, symmetric border.
extern "C" void or_gauss(const double* in, int H, int W, double sigma, double* out) {
  const int r = (int)std::ceil(3.0 * sigma);
  std::vector<double> k(2 * r + 1);
  double s = 0.0;
  for (int i = -r; i <= r; ++i) {
    k[i + r] = std::exp(-(double)(i * i) / (2.0 * sigma * sigma));
    s += k[i + r];
  }
  for (auto& v : k) v /= s;
  std::vector<double> tmp((long)H * W);
  for (int y = 0; y < H; ++y)
    for (int x = 0; x < W; ++x) {
      double a = 0.0;
      for (int i = -r; i <= r; ++i) a += k[i + r] * in[(long)y * W + or_reflect(x + i, W)];
      tmp[(long)y * W + x] = a;
    }
  for (int y = 0; y < H; ++y)
    for (int x = 0; x < W; ++x) {
      double a = 0.0;
      for (int i = -r; i <= r; ++i) a += k[i + r] * tmp[(long)or_reflect(y + i, H) * W + x];
      out[(long)y * W + x] = a;
    }
}

static inline double edge_fn(double d, double alpha, double beta, double lambda) {
  return alpha > 0.0 ? std::max(lambda * std::exp(-alpha * std::pow(d, beta)), 1e-6) : lambda * d;
}

// stereo/stereo.cu:1090-1095 (calcEdgeImage) with the forward-difference
// kernels at stereo/stereo.cu:471-496 (x) and 525-550 (y).
extern "C" void or_edge_weights(const double* gray, int H, int W, double sigma, double alpha,
                                double beta, double lambda, double* wx, double* wy) {
  std::vector<double> g((long)H * W);
  or_gauss(gray, H, W, sigma, g.data());
  for (int y = 0; y < H; ++y)
    for (int x = 0; x < W; ++x) {
      const double c = g[(long)y * W + x];
      wx[(long)y * W + x] = edge_fn(std::fabs(g[(long)y * W + or_reflect(x + 1, W)] - c), alpha, beta, lambda);
      wy[(long)y * W + x] = edge_fn(std::fabs(g[(long)or_reflect(y + 1, H) * W + x] - c), alpha, beta, lambda);
    }
}

// stereo/stereo.cu:348-370: first strict minimum, ties keep the lowest index.
extern "C" void or_argmin(const double* vol, long HW, int D, int* idx) {
  for (long p = 0; p < HW; ++p) {
    double best = std::numeric_limits<double>::max();
    int bi = 0;
    for (int k = 0; k < D; ++k)
      if (vol[p * D + k] < best) {
        best = vol[p * D + k];
        bi = k;
      }
    idx[p] = bi;
  }
}

// stereo/stereo.cu:841-850.
extern "C" void or_remap(const int* labels, long HW, double disp_min, double disp_step, double* out) {
  for (long p = 0; p < HW; ++p) out[p] = (double)labels[p] * disp_step + disp_min;
}

// QuadDirect refinement, stereo/stereo.cu:924-942 (do_fitting_kernel_volume,
// lambda = 1) and :944-953 (do_argmin_kernel_volume): z = (int)((d - min) /
// step); for 0 < z < D - 1, d += g / k with g = (c[z-1] - c[z+1]) / (2 step)
// and k = max(1e-3, (c[z-1] + c[z+1] - 2 c[z]) / step^2).  The reference skips
// z = 0 and reads past the pixel at z = D - 1; both keep d here.
extern "C" void or_quad_fit(double* disp, const double* vol, long HW, int D, double disp_min, double disp_step) {
  for (long p = 0; p < HW; ++p) {
    const int z = (int)((disp[p] - disp_min) / disp_step);
    if (z <= 0 || z >= D - 1) continue;
    const double* c = vol + p * D + z;
    const double g = (c[-1] - c[1]) / (2.0 * disp_step);
    const double k = std::max(1e-3, (c[-1] + c[1] - 2.0 * c[0]) / (disp_step * disp_step));
    disp[p] += g / k;
  }
}

// Left-right check, stereo/stereo.cu:803-817 with the clamped linear texture of
// :84-104: R sampled at texel coordinate x + 0.5 - I1 (linear weights, borders
// clamped); out = I1 where |I1 + R| <= th, else -1.
extern "C" void or_lr_check(const double* left, const double* right, int H, int W, double th, double* out) {
  for (int y = 0; y < H; ++y)
    for (int x = 0; x < W; ++x) {
      const double I1 = left[(long)y * W + x];
      const double u = x - I1;  // (x + 0.5 - I1) - 0.5
      const double fi = std::floor(u);
      const double a = u - fi;
      const int i0 = (int)std::min(std::max(fi, 0.0), (double)(W - 1));
      const int i1 = (int)std::min(std::max(fi + 1.0, 0.0), (double)(W - 1));
      const double I2 = -((1.0 - a) * right[(long)y * W + i0] + a * right[(long)y * W + i1]);
      out[(long)y * W + x] = std::fabs(I1 - I2) <= th ? I1 : -1.0;
    }
}

extern "C" double or_rho(int d, double L1, double L2, int delta) {
  const int a = d < 0 ? -d : d;
  if (a == 0) return 0.0;
  return a < delta ? L1 : L2;
}

extern "C" double or_energy(const double* C, const double* wx, const double* wy, int H, int W, int D,
                            const int* labels, double L1, double L2, int delta) {
  double e = 0.0;
  for (int y = 0; y < H; ++y)
    for (int x = 0; x < W; ++x) {
      const long p = (long)y * W + x;
      const int l = labels[p];
      e += C[p * D + l];
      if (x + 1 < W) e += wx[p] * or_rho(l - labels[p + 1], L1, L2, delta);
      if (y + 1 < H) e += wy[p] * or_rho(l - labels[p + W], L1, L2, delta);
    }
  return e;
}

// The energy over the labelled part of the grid: pixels whose mask label lies
// outside [0, D) drop out with their unary and every edge they end.  The
// structured-SVM energies of training with sparse ground truth (the reference's
// training code is not in the repository; this is the definition the GPU's
// masked energy and subgradient follow).
extern "C" double or_energy_masked(const double* C, const double* wx, const double* wy, int H, int W, int D,
                                   const int* labels, const int* mask, double L1, double L2, int delta) {
  auto in = [&](long q) { return mask[q] >= 0 && mask[q] < D; };
  double e = 0.0;
  for (int y = 0; y < H; ++y)
    for (int x = 0; x < W; ++x) {
      const long p = (long)y * W + x;
      if (!in(p)) continue;
      const int l = labels[p];
      e += C[p * D + l];
      if (x + 1 < W && in(p + 1)) e += wx[p] * or_rho(l - labels[p + 1], L1, L2, delta);
      if (y + 1 < H && in(p + W)) e += wy[p] * or_rho(l - labels[p + W], L1, L2, delta);
    }
  return e;
}

namespace {

// The chains below take each edge's penalties P1 (0 < |d| < delta) and P2
// (|d| >= delta): w L1 and w L2 for the edge-weight form, per-edge values for
// the pairwise-volume form.
//
// out(s) = min_t U(t) + rho_e(s - t), exact for 0 <= P1 <= P2:
// min(U(s), min_{0<|s-t|<delta} U(t) + P1, min_t U(t) + P2).
void minconv(const double* U, int D, double P1, double P2, int delta, double* out) {
  double g = kInf;
  for (int t = 0; t < D; ++t) g = std::min(g, U[t]);
  const double far = g + P2;
  for (int s = 0; s < D; ++s) {
    double best = U[s];
    double near = kInf;
    for (int t = std::max(0, s - delta + 1); t <= std::min(D - 1, s + delta - 1); ++t)
      if (t != s) near = std::min(near, U[t]);
    best = std::min(best, near + P1);
    best = std::min(best, far);
    out[s] = best;
  }
}

double vmin(const double* v, int D) {
  double m = kInf;
  for (int k = 0; k < D; ++k) m = std::min(m, v[k]);
  return m;
}

// Backward pass shared by the minorant and the decode:
// Bh[n-1] = 0; U = theta[i+1] + Bh[i+1]; u_i = min U; Bh[i] = minconv(U) - u_i.
// The true backward message is B_i = Bh[i] + b_i with b_{n-1} = 0,
// b_i = b_{i+1} + u_i; b (n entries) receives these offsets.
void backward_pass(const double* theta, const double* p1, const double* p2, int n, int D, int delta, double* Bh,
                   double* b) {
  std::vector<double> U(D);
  for (int k = 0; k < D; ++k) Bh[(long)(n - 1) * D + k] = 0.0;
  b[n - 1] = 0.0;
  for (int i = n - 2; i >= 0; --i) {
    for (int k = 0; k < D; ++k) U[k] = theta[(long)(i + 1) * D + k] + Bh[(long)(i + 1) * D + k];
    const double u = vmin(U.data(), D);
    b[i] = b[i + 1] + u;
    minconv(U.data(), D, p1[i], p2[i], delta, Bh + (long)i * D);
    for (int k = 0; k < D; ++k) Bh[(long)i * D + k] -= u;
  }
}

// excess(s) = min_t [rho_e(s - t) - m(t)] + V(s) for the message m = minconv(V):
// the part of V(s) that the message to the next node cannot see.  Because
// m(t) >= min V >= max m - P2, the far term never binds, so
// excess(s) = V(s) - max(m(s), max_{0<|s-t|<delta} m(t) - P1)  (>= 0, min 0).
void excess(const double* V, const double* m, int D, double P1, int delta, double* e) {
  for (int s = 0; s < D; ++s) {
    double near = -kInf;
    for (int t = std::max(0, s - delta + 1); t <= std::min(D - 1, s + delta - 1); ++t)
      if (t != s) near = std::max(near, m[t]);
    e[s] = V[s] - std::max(m[s], near - P1);
  }
}

// Tight, maximal modular minorant of a chain (slack split gamma = 1/4, last node 1):
//   g_i = gamma (R_i + B_i) + excess_i,   V_i = R_i - gamma (R_i + B_i),
//   R_{i+1} = theta_{i+1} + minconv(V_i),   g_{n-1} = R_{n-1}.
// sum_i g_i(y_i) <= E(y) for every y, sum_i min g_i = min E (tightness), and the
// remaining pairwise slack is arc consistent in both directions.
// g is stored without its per-node constant kappa_i (g_i = g~_i + kappa_i); the
// constants follow the scalar recurrence below and *K receives sum_i kappa_i,
// which is exactly the chain minimum left behind when g~ is subtracted.
double minorant_impl(const double* theta, const double* p1, const double* p2, int n, int D, int delta, double* g,
                     double* K, std::vector<double>& Bh) {
  constexpr double gamma = 0.25;
  Bh.resize((long)n * D);
  std::vector<double> b(n);
  backward_pass(theta, p1, p2, n, D, delta, Bh.data(), b.data());
  double m0 = kInf;
  for (int k = 0; k < D; ++k) m0 = std::min(m0, theta[k] + Bh[k]);
  const double Lambda = b[0] + m0;
  std::vector<double> R(theta, theta + D), V(D), msg(D), e(D);
  double r = 0.0, kept = 0.0;  // R_i = R~_i + r
  for (int i = 0; i + 1 < n; ++i) {
    double* gi = g + (long)i * D;
    for (int k = 0; k < D; ++k) {
      gi[k] = gamma * (R[k] + Bh[(long)i * D + k]);
      V[k] = R[k] - gi[k];
    }
    kept += gamma * (r + b[i]);
    const double v = vmin(V.data(), D);
    minconv(V.data(), D, p1[i], p2[i], delta, msg.data());
    excess(V.data(), msg.data(), D, p1[i], delta, e.data());
    for (int k = 0; k < D; ++k) {
      gi[k] += e[k];
      R[k] = theta[(long)(i + 1) * D + k] + msg[k] - v;
    }
    r = v + (1.0 - gamma) * r - gamma * b[i];
  }
  for (int k = 0; k < D; ++k) g[(long)(n - 1) * D + k] = R[k];
  kept += r;
  if (K) *K = kept;
  return Lambda;
}

// Decode of one chain: labels = argmin of the min-marginals F_i + B_i
// (lowest label among those within kTieEps of the minimum); gap = the best
// value over the other labels minus the minimum (0 when the chosen label is
// not the exact minimum).
double decode_impl(const double* theta, const double* p1, const double* p2, int n, int D, int delta, int* labels,
                   double* gap, std::vector<double>& Bh) {
  Bh.resize((long)n * D);
  std::vector<double> b(n);
  backward_pass(theta, p1, p2, n, D, delta, Bh.data(), b.data());
  double m0 = kInf;
  for (int k = 0; k < D; ++k) m0 = std::min(m0, theta[k] + Bh[k]);
  std::vector<double> F(theta, theta + D), msg(D);
  for (int i = 0; i < n; ++i) {
    if (i > 0) {
      const double f = vmin(F.data(), D);
      minconv(F.data(), D, p1[i - 1], p2[i - 1], delta, msg.data());
      for (int k = 0; k < D; ++k) F[k] = theta[(long)i * D + k] + msg[k] - f;
    }
    // Lowest label whose min-marginal is within kTieEps (relative) of the
    // minimum: min-marginals on a plateau of the truncated pairwise term tie in
    // exact arithmetic but differ by rounding, and the lowest-label rule must
    // see them as ties (a spatially coherent choice).
    double best = kInf;
    for (int k = 0; k < D; ++k) best = std::min(best, F[k] + Bh[(long)i * D + k]);
    const double thr = best + kTieEps * std::max(1.0, std::fabs(best));
    int bl = 0;
    while (bl < D - 1 && !(F[bl] + Bh[(long)i * D + bl] <= thr)) ++bl;
    double second = kInf;
    for (int k = 0; k < D; ++k)
      if (k != bl) second = std::min(second, F[k] + Bh[(long)i * D + k]);
    labels[i] = bl;
    if (gap) gap[i] = second - best;
  }
  return b[0] + m0;
}

// Per-edge penalties of a chain of n nodes from edge weights: P1 = w L1, P2 = w L2.
void weighted(const double* w, int n, double L1, double L2, std::vector<double>& p1, std::vector<double>& p2) {
  p1.assign(n > 1 ? n - 1 : 0, 0.0);
  p2.assign(p1.size(), 0.0);
  for (int i = 0; i + 1 < n; ++i) {
    p1[i] = w[i] * L1;
    p2[i] = w[i] * L2;
  }
}

// The restated SlackProp on per-edge penalties: rows use the horizontal
// edges' (p1x, p2x)[y][x] (edge (x,y)-(x+1,y)), columns the vertical edges'
// (p1y, p2y)[y][x] (edge (x,y)-(x,y+1)).  See or_slackprop.
double slackprop_impl(const double* C, const double* p1x, const double* p2x, const double* p1y, const double* p2y,
                      int H, int W, int D, int delta, int iters, int* labels, double* sweep_sums, double* kept_sums,
                      double* gap, double* A_out) {
  const long HW = (long)H * W;
  std::vector<double> A(C, C + HW * D);
  std::vector<double> sums(2 * (long)iters + 1, 0.0), kept(2 * (long)iters, 0.0);
  int j = 0;
  for (int it = 0; it < iters; ++it) {
    double srow = 0.0, krow = 0.0;
#pragma omp parallel for schedule(dynamic) reduction(+ : srow, krow)
    for (int y = 0; y < H; ++y) {
      std::vector<double> g((long)W * D), Bh;
      double* Ar = A.data() + (long)y * W * D;
      double k = 0.0;
      srow += minorant_impl(Ar, p1x + (long)y * W, p2x + (long)y * W, W, D, delta, g.data(), &k, Bh);
      krow += k;
      for (long i = 0; i < (long)W * D; ++i) Ar[i] -= g[i];
    }
    sums[j] = srow;
    kept[j] = krow;
    ++j;
    double scol = 0.0, kcol = 0.0;
#pragma omp parallel for schedule(dynamic) reduction(+ : scol, kcol)
    for (int x = 0; x < W; ++x) {
      std::vector<double> th((long)H * D), g((long)H * D), c1(H), c2(H), Bh;
      for (int y = 0; y < H; ++y) {
        c1[y] = p1y[(long)y * W + x];
        c2[y] = p2y[(long)y * W + x];
        for (int k = 0; k < D; ++k) th[(long)y * D + k] = C[((long)y * W + x) * D + k] - A[((long)y * W + x) * D + k];
      }
      double k = 0.0;
      scol += minorant_impl(th.data(), c1.data(), c2.data(), H, D, delta, g.data(), &k, Bh);
      kcol += k;
      for (int y = 0; y < H; ++y)
        for (int k = 0; k < D; ++k) A[((long)y * W + x) * D + k] += g[(long)y * D + k];
    }
    sums[j] = scol;
    kept[j] = kcol;
    ++j;
  }
  double sdec = 0.0;
#pragma omp parallel for schedule(dynamic) reduction(+ : sdec)
  for (int y = 0; y < H; ++y) {
    std::vector<double> Bh;
    sdec += decode_impl(A.data() + (long)y * W * D, p1x + (long)y * W, p2x + (long)y * W, W, D, delta,
                        labels + (long)y * W, gap ? gap + (long)y * W : nullptr, Bh);
  }
  sums[j] = sdec;
  if (sweep_sums) std::memcpy(sweep_sums, sums.data(), sizeof(double) * sums.size());
  if (kept_sums && iters > 0) std::memcpy(kept_sums, kept.data(), sizeof(double) * kept.size());
  if (A_out) std::memcpy(A_out, A.data(), sizeof(double) * HW * D);
  return iters == 0 ? sdec : kept[j - 1] + sdec;
}

}  // namespace

extern "C" double or_chain_viterbi(const double* theta, const double* w, int n, int D, double L1,
                                   double L2, int delta, int* labels) {
  std::vector<double> F((long)n * D);
  std::vector<int> arg((long)n * D);
  for (int k = 0; k < D; ++k) F[k] = theta[k];
  for (int i = 1; i < n; ++i)
    for (int t = 0; t < D; ++t) {
      double best = kInf;
      int ba = 0;
      for (int s = 0; s < D; ++s) {
        const double v = F[(long)(i - 1) * D + s] + w[i - 1] * or_rho(s - t, L1, L2, delta);
        if (v < best) {
          best = v;
          ba = s;
        }
      }
      F[(long)i * D + t] = theta[(long)i * D + t] + best;
      arg[(long)i * D + t] = ba;
    }
  double best = kInf;
  int bl = 0;
  for (int t = 0; t < D; ++t)
    if (F[(long)(n - 1) * D + t] < best) {
      best = F[(long)(n - 1) * D + t];
      bl = t;
    }
  if (labels) {
    labels[n - 1] = bl;
    for (int i = n - 1; i > 0; --i) labels[i - 1] = arg[(long)i * D + labels[i]];
  }
  return best;
}

extern "C" double or_chain_minorant(const double* theta, const double* w, int n, int D, double L1,
                                    double L2, int delta, double* g, double* kept) {
  std::vector<double> Bh, p1, p2;
  weighted(w, n, L1, L2, p1, p2);
  return minorant_impl(theta, p1.data(), p2.data(), n, D, delta, g, kept, Bh);
}

// Restated SlackProp.  State: row unaries A (the columns hold C - A).
//   A = C
//   repeat iters times:
//     rows:    g~ = minorant(A_row, wx_row);       A_row -= g~
//     columns: g~ = minorant((C-A)_col, wy_col);   A_col += g~
//   decode rows of A: labels = argmin of row min-marginals
// Half-step j (1-based) records S_j = sum of the swept chains' minima and
// K_j = sum of the constants kappa the swept chains keep.  Each half-step moves
// a tight minorant of one chain family into the other, so the dual
// Phi_rows(A) + Phi_cols(C - A) never decreases: it is S_1 before the first
// half-step and K_j + S_{j+1} after half-step j.  Edge weights: wx[y][x]
// couples (x,y)-(x+1,y) and wy[y][x] couples (x,y)-(x,y+1), as produced by
// stereo/stereo.cu:471-550.
extern "C" double or_slackprop(const double* C, const double* wx, const double* wy, int H, int W,
                               int D, double L1, double L2, int delta, int iters, int* labels,
                               double* sweep_sums, double* kept_sums, double* gap, double* A_out) {
  const long HW = (long)H * W;
  std::vector<double> p1x(HW), p2x(HW), p1y(HW), p2y(HW);
  for (long p = 0; p < HW; ++p) {
    p1x[p] = wx[p] * L1;
    p2x[p] = wx[p] * L2;
    p1y[p] = wy[p] * L1;
    p2y[p] = wy[p] * L2;
  }
  return slackprop_impl(C, p1x.data(), p2x.data(), p1y.data(), p2y.data(), H, W, D, delta, iters, labels, sweep_sums,
                        kept_sums, gap, A_out);
}

// The pairwise-volume form of SlackProp::forward (stereo/operators/slackprop.cpp:71-108):
// per-edge penalties instead of weighted L1, L2.  The state, schedule, bound
// and decode are those of or_slackprop.
extern "C" double or_slackprop_pe(const double* C, const double* p1x, const double* p2x, const double* p1y,
                                  const double* p2y, int H, int W, int D, int delta, int iters, int* labels,
                                  double* sweep_sums, double* kept_sums, double* gap, double* A_out) {
  return slackprop_impl(C, p1x, p2x, p1y, p2y, H, W, D, delta, iters, labels, sweep_sums, kept_sums, gap, A_out);
}

// E(x) with per-edge penalties: rho_e(d) = 0, P1 (0 < |d| < delta), P2 (|d| >= delta).
extern "C" double or_energy_pe(const double* C, const double* p1x, const double* p2x, const double* p1y,
                               const double* p2y, int H, int W, int D, const int* labels, int delta) {
  auto r = [delta](int d, double P1, double P2) {
    const int a = d < 0 ? -d : d;
    return a == 0 ? 0.0 : (a < delta ? P1 : P2);
  };
  double e = 0.0;
  for (int y = 0; y < H; ++y)
    for (int x = 0; x < W; ++x) {
      const long p = (long)y * W + x;
      const int l = labels[p];
      e += C[p * D + l];
      if (x + 1 < W) e += r(l - labels[p + 1], p1x[p], p2x[p]);
      if (y + 1 < H) e += r(l - labels[p + W], p1y[p], p2y[p]);
    }
  return e;
}

// The reference's re-layout of the pairwise volume (stereo/operators/slackprop.cpp:88-92):
// pen[l][y][x][c] = pw[c + 2 (levels - 1 - l)][y][x] for pw of 2 levels planes.
extern "C" void or_pw_relayout(const double* pw, int levels, int H, int W, double* pen) {
  const long HW = (long)H * W;
  for (int l = 0; l < levels; ++l)
    for (long p = 0; p < HW; ++p)
      for (int c = 0; c < 2; ++c) pen[((long)l * HW + p) * 2 + c] = pw[(long)(c + 2 * (levels - 1 - l)) * HW + p];
}

extern "C" void or_ssvm_grad(const int* xhat, const int* xstar, const double* wx, const double* wy,
                             int H, int W, int D, double L1, double L2, int delta, double* dC,
                             double* dL, double* dwx, double* dwy) {
  (void)L1;
  (void)L2;
  const long HW = (long)H * W;
  // Pixels whose x* lies outside [0, D) are unlabelled: both energies leave
  // out their unary and their edges (or_energy_masked with mask = x*).
  auto lab = [&](long q) { return xstar[q] >= 0 && xstar[q] < D; };
  if (dC) {
    std::memset(dC, 0, sizeof(double) * HW * D);
    for (long p = 0; p < HW; ++p) {
      if (!lab(p)) continue;
      dC[p * D + xstar[p]] += 1.0;
      dC[p * D + xhat[p]] -= 1.0;
    }
  }
  double d1 = 0.0, d2 = 0.0;
  auto level = [delta](int a, int b, int which) {
    const int d = a > b ? a - b : b - a;
    if (d == 0) return 0.0;
    return (which == 1) == (d < delta) ? 1.0 : 0.0;
  };
  for (int y = 0; y < H; ++y)
    for (int x = 0; x < W; ++x) {
      const long p = (long)y * W + x;
      double gx = 0.0, gy = 0.0;
      if (x + 1 < W && lab(p) && lab(p + 1)) {
        d1 += wx[p] * (level(xstar[p], xstar[p + 1], 1) - level(xhat[p], xhat[p + 1], 1));
        d2 += wx[p] * (level(xstar[p], xstar[p + 1], 2) - level(xhat[p], xhat[p + 1], 2));
        gx = or_rho(xstar[p] - xstar[p + 1], L1, L2, delta) - or_rho(xhat[p] - xhat[p + 1], L1, L2, delta);
      }
      if (y + 1 < H && lab(p) && lab(p + W)) {
        d1 += wy[p] * (level(xstar[p], xstar[p + W], 1) - level(xhat[p], xhat[p + W], 1));
        d2 += wy[p] * (level(xstar[p], xstar[p + W], 2) - level(xhat[p], xhat[p + W], 2));
        gy = or_rho(xstar[p] - xstar[p + W], L1, L2, delta) - or_rho(xhat[p] - xhat[p + W], L1, L2, delta);
      }
      if (dwx) dwx[p] = gx;
      if (dwy) dwy[p] = gy;
    }
  if (dL) {
    dL[0] = d1;
    dL[1] = d2;
  }
}
