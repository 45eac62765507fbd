/*
 * FP64 CPU oracle of the CNN-CRF stereo hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * product path is compared against.  Only tests, the smoke check and the
 * CPU-baseline leg of the benchmark may load it; nothing in the product
 * links, calls or falls back to it.
 *
 * Every function restates one operator of the reference path in double
 * precision, in plain loops, and cites the reference file:line it follows.
 * Tensors are dense row-major double arrays:
 *   images / feature maps   C x H x W   (the reference's NCHW with N = 1)
 *   cost volumes            H x W x D   (the reference's NHWC, disparity fastest)
 *   edge weights            H x W
 *   labels                  H x W int32
 */
#ifndef CCS_ORACLE_H
#define CCS_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Symmetric (edge-repeating) reflection of index i into [0, n). */
int or_reflect(int i, int n);

/* Mirror padding: stereo/stereo.cu:865-889 (texture cudaAddressModeMirror). */
void or_pad_symmetric(const double* in, int C, int H, int W, int p, double* out);

/* Per-channel zero mean / unit population std: stereo/utils.cu:169-237, 242-287. */
void or_normalize(double* x, int C, long HW, double* mean_out, double* std_out);

/* rgb -> gray: stereo/stereo.cu:853-863. rgb is 3 x H x W. */
void or_rgb2gray(const double* rgb, int H, int W, double* gray);

/* Valid true convolution (kernel flipped) + bias + optional tanh:
 * stereo/operators/convolution.cpp:104-152, bias.cpp:58-66, activation.cpp:33-40.
 * x: C x H x W, w: K x C x KH x KW, b: K (may be NULL), y: K x (H-KH+1) x (W-KW+1). */
void or_conv2d(const double* x, int C, int H, int W, const double* w, const double* b,
               int K, int KH, int KW, int act_tanh, double* y);

/* Adjoint of or_conv2d (activation recomputed from x, w, b).
 * dy is the gradient w.r.t. the output after the activation.
 * dx (C x H x W), dw (K x C x KH x KW), db (K) are overwritten; any may be NULL. */
void or_conv2d_bwd(const double* x, int C, int H, int W, const double* w, const double* b,
                   int K, int KH, int KW, int act_tanh, const double* dy,
                   double* dx, double* dw, double* db);

/* Correlation: stereo/operators/stereocorrelation.cu:111-142.
 * out[h][w][k] = max_{|y|<=rc, 0<=h+y<H} ( 0<=w+disp[k]<W ? sum_c f0[c][h][w] f1[c][h+y][w+disp[k]] : 0 ).
 * disp[] holds the reference's disparity offsets (-d for left->right matching). */
void or_correlation(const double* f0, const double* f1, int C, int H, int W,
                    const int* disp, int D, int rect_corr, double* out);

/* Adjoint of or_correlation for rect_corr == 0:
 * stereo/operators/stereocorrelation.cu:243-278. G is H x W x D. */
void or_correlation_bwd(const double* f0, const double* f1, int C, int H, int W,
                        const int* disp, int D, const double* G, double* df0, double* df1);

/* cost = -softmax_D(vol), in place: stereo/operators/softmax.cpp:30-38, stereo/utils.cu:289-292. */
void or_neg_softmax(double* vol, long HW, int D);

/* Cross entropy of softmax_D(S) at target labels (targets outside [0, D) ignored).
 * Returns the summed loss; dS (H*W*D) receives dLoss/dS. */
double or_softmax_xent(const double* S, long HW, int D, const int* target, double* dS);

/* Gaussian smoothing, radius ceil(3 sigma), symmetric border (stands in for iu::filterGauss). */
void or_gauss(const double* in, int H, int W, double sigma, double* out);

/* Edge weights: stereo/stereo.cu:471-496, 525-550, 1090-1095.
 * g = gauss(gray, sigma); wx = f(|g(x+1,y)-g(x,y)|), wy = f(|g(x,y+1)-g(x,y)|), mirrored border,
 * f(d) = alpha > 0 ? max(lambda exp(-alpha d^beta), 1e-6) : lambda d. */
void or_edge_weights(const double* gray, int H, int W, double sigma, double alpha, double beta,
                     double lambda, double* wx, double* wy);

/* First strict minimum over D: stereo/stereo.cu:348-370. */
void or_argmin(const double* vol, long HW, int D, int* idx);

/* label -> disparity: stereo/stereo.cu:841-850. */
void or_remap(const int* labels, long HW, double disp_min, double disp_step, double* out);

/* QuadDirect sub-pixel refinement of disp in place (stereo/stereo.cu:924-953). */
void or_quad_fit(double* disp, const double* vol, long HW, int D, double disp_min, double disp_step);

/* Left-right consistency check (stereo/stereo.cu:803-817). */
void or_lr_check(const double* left, const double* right, int H, int W, double th, double* out);

/* Pairwise penalty rho of the CRF (0 / L1 for 0<|d|<delta / L2 for |d|>=delta). */
double or_rho(int d, double L1, double L2, int delta);

/* CRF energy E(x) = sum_p C_p(x_p) + sum_h-edges wx rho + sum_v-edges wy rho. */
double or_energy(const double* C, const double* wx, const double* wy, int H, int W, int D,
                 const int* labels, double L1, double L2, int delta);

/* or_energy restricted to the pixels whose mask label lies in [0, D): the others
 * drop out with their unary and every edge they end (training with sparse ground truth). */
double or_energy_masked(const double* C, const double* wx, const double* wy, int H, int W, int D,
                        const int* labels, const int* mask, double L1, double L2, int delta);

/* Exact minimum of one chain (Viterbi), theta n x D, w n-1 edge weights.
 * Returns E*, writes an optimal labeling (lowest label on ties, backtracked). */
double or_chain_viterbi(const double* theta, const double* w, int n, int D,
                        double L1, double L2, int delta, int* labels);

/* Tight modular minorant of one chain (the SlackProp minorize step).
 * g (n x D) receives the minorant without its per-node constants kappa_i;
 * *kept (optional) receives sum_i kappa_i.  Returns Lambda = min chain energy. */
double or_chain_minorant(const double* theta, const double* w, int n, int D,
                         double L1, double L2, int delta, double* g, double* kept);

/* Restated SlackProp (dual minorize-maximize over row and column chains).
 * C is H x W x D.  Writes labels (H x W).  sweep_sums (2*iters+1) receives the
 * per-half-step chain minima S_j, kept_sums (2*iters) the constants K_j the swept
 * chains keep; the dual bound is S_1 initially and K_j + S_{j+1} after half-step j.
 * gap (H x W, optional) receives second-best minus best decode min-marginal.
 * A_out (H x W x D, optional) receives the final row unaries.  Returns the final
 * lower bound K_{2 iters} + S_{2 iters + 1}.  Call-site semantics follow
 * stereo/stereo.cpp:140-154 and stereo/operators/slackprop.cpp:111-139. */
double or_slackprop(const double* C, const double* wx, const double* wy, int H, int W, int D,
                    double L1, double L2, int delta, int iters, int* labels,
                    double* sweep_sums, double* kept_sums, double* gap, double* A_out);

/* or_slackprop with per-edge penalties (the pairwise-volume form of SlackProp::forward,
 * stereo/operators/slackprop.cpp:71-108): the horizontal edge (x,y)-(x+1,y) costs
 * p1x[y][x] for 0 < |d| < delta and p2x[y][x] for |d| >= delta, the vertical edge
 * (x,y)-(x,y+1) p1y / p2y.  Same schedule, bound and decode as or_slackprop. */
double or_slackprop_pe(const double* C, const double* p1x, const double* p2x, const double* p1y,
                       const double* p2y, int H, int W, int D, int delta, int iters, int* labels,
                       double* sweep_sums, double* kept_sums, double* gap, double* A_out);

/* The CRF energy with per-edge penalties (the energy or_slackprop_pe bounds). */
double or_energy_pe(const double* C, const double* p1x, const double* p2x, const double* p1y,
                    const double* p2y, int H, int W, int D, const int* labels, int delta);

/* The pairwise volume's re-layout (stereo/operators/slackprop.cpp:88-92):
 * pen[l][y][x][c] = pw[c + 2 (levels - 1 - l)][y][x]. */
void or_pw_relayout(const double* pw, int levels, int H, int W, double* pen);

/* Structured-SVM subgradient of  E(x*) - E(xhat)  w.r.t. unaries, L1, L2 and edge weights.
 * dC (H x W x D) gets +1 at x*, -1 at xhat; dL[0], dL[1] get dE/dL1, dE/dL2 differences;
 * dwx, dwy (H x W, optional) get rho differences.  Pixels with x* outside [0, D) are
 * unlabelled and drop out of both energies (or_energy_masked with mask = x*). */
void or_ssvm_grad(const int* xhat, const int* xstar, const double* wx, const double* wy,
                  int H, int W, int D, double L1, double L2, int delta,
                  double* dC, double* dL, double* dwx, double* dwy);

#ifdef __cplusplus
}
#endif
#endif
