/*
 * ccs.h -- C ABI of the MI355X-native CNN-CRF stereo hot path.
 *
 * Plain C: pointers, sizes and status codes only.  Every device pointer is
 * caller-owned HBM; work is enqueued on the handle's HIP stream and nothing
 * blocks unless a function says so.  A handle's scratch memory grows on first
 * use at a size, in the order of its stream (no host wait); a call that would
 * grow it while the stream is being captured into a HIP graph returns
 * CCS_ERR_ARG instead, so warm a handle up (ccs_reserve) before capturing.
 * Functions return CCS_OK (0) or a negative status; ccs_last_error() gives the
 * message of the calling thread's last failure.  Each entry point names the reference interface it replaces
 * (paths relative to the VLOGroup/cnn-crf-stereo source tree).
 *
 * Layouts
 *   image        C x H x W float32                      (reference NCHW, N = 1)
 *   feature map  H x W x Cs, Cs = ccs_feature_stride(C) (channel-last, channels
 *                C..Cs-1 hold zeros), element type float32 or bfloat16
 *   cost volume  H x W x D float32, disparity fastest   (reference NHWC)
 *   edge weight  H x W float32; wx[y][x] couples (x,y)-(x+1,y), wy[y][x] couples (x,y)-(x,y+1)
 *   labels       H x W int32
 */
#ifndef CCS_H
#define CCS_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ccs_handle_s* ccs_handle_t;

enum {
  CCS_OK = 0,
  CCS_ERR_ARG = -1,         /* invalid argument or size                       */
  CCS_ERR_UNSUPPORTED = -2, /* valid in the reference but not offered here     */
  CCS_ERR_HIP = -3,         /* HIP runtime error                                */
  CCS_ERR_ALLOC = -4,       /* device allocation failed                         */
  CCS_ERR_IO = -5,          /* file missing or malformed                        */
  CCS_ERR_SHAPE = -6,       /* tensor shape does not match the operator         */
  CCS_ERR_NAN = -7          /* an input holds NaN where the operator needs numbers */
};

enum { CCS_LAYOUT_NCHW = 0, CCS_LAYOUT_NHWC = 1 };    /* iu::TensorGpu_32f::MemoryLayout */
enum { CCS_FP32 = 0, CCS_BF16 = 1 };                  /* arithmetic of conv / correlation */
enum { CCS_EPILOGUE_RAW = 0, CCS_EPILOGUE_NEG_SOFTMAX = 1 };

const char* ccs_last_error(void);
const char* ccs_version(void);

/* Handle: a HIP stream plus grow-only scratch memory owned by the handle.
 * Replaces cudnnCreate / cudnnDestroy (stereo/stereonet.cpp:38,43).
 * stream == NULL uses the null stream. */
int ccs_create(ccs_handle_t* handle, void* hip_stream);
int ccs_destroy(ccs_handle_t handle);
/* Switching streams waits for the work queued on the old one (blocking). */
int ccs_set_stream(ccs_handle_t handle, void* hip_stream);
int ccs_synchronize(ccs_handle_t handle);

/* Grows the handle's scratch to what the forward entry points of this header
 * need for an H x W frame with D labels, C feature channels (and 64-channel
 * pairwise layers) and `iters` CRF iterations, so that later calls at or below
 * those sizes allocate nothing and may be captured into a HIP graph.  The
 * adjoints (ccs_conv2d_backward) are not covered.  Stream-ordered.  No
 * reference counterpart (cudnnHandle_t workspaces were sized per call). */
int ccs_reserve(ccs_handle_t handle, int H, int W, int D, int C, int iters);

/* Channel stride of a feature map with C channels (C rounded up to a multiple of 8). */
int ccs_feature_stride(int C);

/* ---- input preparation ------------------------------------------------- */

/* Mirror padding + per-channel zero mean / unit population std of one image.
 * Replaces padColorImage / padding_kernel (stereo/stereo.cu:865-889,1022-1044)
 * followed by makeRgbZeroMean + makeRgbUnitStd (stereo/utils.cu:169-237) or
 * makeZeroMean + makeUnitStd (stereo/utils.cu:242-287).
 * img: C x H x W float32; out: C x (H+2p) x (W+2p) float32.
 * stats (optional, device, 2*C doubles) receives mean[0..C) and std[0..C). */
int ccs_pad_normalize(ccs_handle_t h, const float* img, int C, int H, int W, int pad,
                      float* out, double* stats);

/* 0.2989 R + 0.5870 G + 0.1140 B of a 3 x H x W image: stereo/stereo.cu:853-863. */
int ccs_rgb_to_gray(ccs_handle_t h, const float* rgb, int H, int W, float* gray);

/* ---- feature CNN --------------------------------------------------------- */

/* Fused Convolution + Bias + Activation(tanh) of one tower layer.  Replaces
 * Convolution::forward (stereo/operators/convolution.cpp:146-152, CUDNN_CONVOLUTION,
 * i.e. the filter is flipped, pad 0, stride 1), Bias::forward
 * (stereo/operators/bias.cpp:58-66) and Activation::forward
 * (stereo/operators/activation.cpp:33-40).
 * x: H x W x ccs_feature_stride(C) (or, with x_layout == CCS_LAYOUT_NCHW and
 *    C <= 8, a C x H x W float32 image); y: (H-KH+1) x (W-KW+1) x ccs_feature_stride(K).
 * w: K x C x KH x KW float32 (the reference's npz layout), b: K float32 (or NULL).
 * precision CCS_BF16: x, y are bfloat16 and the MFMA accumulates in float32.
 * act: 1 = tanh, 0 = identity.  Weights are repacked into handle scratch on
 * every call; ccs_conv_pack_weights + ccs_conv_forward_packed avoid that. */
int ccs_conv2d_forward(ccs_handle_t h, const void* x, int x_layout, int H, int W, int C,
                       const float* w, const float* b, int K, int KH, int KW, int act,
                       int precision, void* y);

/* ccs_conv2d_forward without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The weights
 * are packed into one fixed 512 KiB array per GPU, held until the process
 * ends, instead of the handle's scratch; every MFMA shape fits (100 -> 100 3x3
 * in FP32 packs to 419,328 bytes) and a larger pack returns
 * CCS_ERR_UNSUPPORTED.  All calls on one GPU share that array, which keeps the
 * last call's packed weights: calls that overlap on one GPU (other streams,
 * other host threads) race.  Same arguments, outputs and kernels as
 * ccs_conv2d_forward. */
int ccs_conv2d_forward_noalloc(ccs_handle_t h, const void* x, int x_layout, int H, int W, int C,
                               const float* w, const float* b, int K, int KH, int KW, int act,
                               int precision, void* y);

size_t ccs_conv_packed_bytes(int K, int C, int KH, int KW, int precision);
int ccs_conv_pack_weights(ccs_handle_t h, const float* w, int K, int C, int KH, int KW,
                          int precision, void* packed);
int ccs_conv_forward_packed(ccs_handle_t h, const void* x, int x_layout, int H, int W, int C,
                            const void* packed, const float* b, int K, int KH, int KW, int act,
                            int precision, void* y);

/* Adjoint of ccs_conv2d_forward (float32 feature maps, tanh recomputed from x).
 * dy: (H-KH+1) x (W-KW+1) x stride(K) gradient w.r.t. the activated output.
 * dx: H x W x stride(C) (may be NULL), dw: K x C x KH x KW, db: K (may be NULL).
 * No reference counterpart (the reference has no convolution backward). */
int ccs_conv2d_backward(ccs_handle_t h, const float* x, int H, int W, int C, const float* w,
                        const float* b, int K, int KH, int KW, int act, const float* dy,
                        float* dx, float* dw, float* db);

/* Adjoint of one layer from its kept (or recomputed) input x (H x W x stride(C),
 * element type x_precision) and activated output y ((H-KH+1) x (W-KW+1) x
 * stride(K), element type precision: CCS_FP32 float32, CCS_BF16 bfloat16);
 * tanh' is 1 - y^2, so nothing is recomputed.  dy (y's shape) and dx (x's
 * shape, may be NULL) have y's element type; dw (K x C x KH x KW) and db (K,
 * may be NULL) are float32.  x_precision may be CCS_FP32 under a CCS_BF16
 * layer (the first layer reads the float32 image).  The weight gradient runs on
 * v_mfma_f32_16x16x4_f32, dx on the forward MFMA kernels, with fixed-order pixel
 * reductions.  No reference counterpart (the reference has no convolution
 * backward). */
int ccs_conv2d_backward_stored(ccs_handle_t h, const void* x, int x_precision, const void* y,
                               int precision, int H, int W, int C, const float* w, int K, int KH,
                               int KW, int act, const void* dy, void* dx, float* dw, float* db);

/* ---- correlation cost volume -------------------------------------------- */

/* StereoCorrelation::forward static (stereo/operators/stereocorrelation.h:38-40,
 * stereocorrelation.cu:386-408): same argument order and meaning.  input0/1 are
 * N x C x H x W float32, disparities is a device array of the reference's
 * integer-valued offsets (-d for left->right), output is N x H x W x D
 * (memory_layout CCS_LAYOUT_NHWC) or N x D x H x W (CCS_LAYOUT_NCHW).
 * Offsets leaving the image give 0.  Like the reference it reads the
 * disparities back to the host, so it synchronises the stream. */
int ccs_correlation_forward(ccs_handle_t h, const float* input0, const float* input1,
                            const float* disparities, float* output, int in, int ic, int ih,
                            int iw, int len_disps, int memory_layout, int rect_corr);

/* StereoCorrelation::backward static (stereo/operators/stereocorrelation.h:42-44,
 * stereocorrelation.cu:437-474): dIn0/1 (N x C x H x W) from the output gradient
 * in_grad (N x H x W x D for CCS_LAYOUT_NHWC).  Gathers by disparity band, no
 * atomics; synchronises like the reference. */
int ccs_correlation_backward(ccs_handle_t h, float* out_grad0, float* out_grad1,
                             const float* in_grad, const float* im0, const float* im1,
                             const float* disparities, int n, int c, int ih, int iw,
                             int num_disps, int memory_layout);

/* Fast path of StereoCorrelation + Softmax + negateTensor
 * (stereo/stereonet.cpp:301-318): banded MFMA correlation of two feature maps
 * (H x W x stride(C), float32 or bfloat16) over the offsets
 * -(disp_min + k disp_step), k < D, with the -softmax over D fused
 * (epilogue CCS_EPILOGUE_NEG_SOFTMAX) or the raw correlation (CCS_EPILOGUE_RAW).
 * out: H x W x D float32.  Stream-ordered, no host synchronisation. */
int ccs_correlation_volume(ccs_handle_t h, const void* f0, const void* f1, int precision,
                           int H, int W, int C, int disp_min, int disp_step, int D,
                           int epilogue, float* out);

/* Adjoint of the fast path: G = dLoss/dS (H x W x D float32, gradient of the raw
 * correlation) -> df0, df1 (H x W x stride(C)) in the feature maps' precision
 * (float32 or bfloat16, accumulated in float32).  Each output gathers its
 * disparity band; no atomics. */
int ccs_correlation_volume_backward(ccs_handle_t h, const void* f0, const void* f1, int precision,
                                    int H, int W, int C, int disp_min, int disp_step, int D,
                                    const float* G, void* df0, void* df1);

/* Cross entropy of softmax_D(S) at target labels (targets outside [0, D) ignored):
 * loss (device, 1 double) and dS = softmax - onehot (H x W x D). */
int ccs_softmax_xent(ccs_handle_t h, const float* S, int HW, int D, const int* target,
                     double* loss, float* dS);

/* ---- CRF ----------------------------------------------------------------- */

/* Edge weights: Gaussian(sigma) of the gray image, then forward differences with
 * a mirrored border, w = alpha > 0 ? max(lambda exp(-alpha |d|^beta), 1e-6) : lambda |d|.
 * Replaces calcEdgeImage (stereo/stereo.cu:1090-1095, 471-496, 525-550). */
int ccs_edge_weights(ccs_handle_t h, const float* gray, int H, int W, float sigma, float alpha,
                     float beta, float lambda, float* wx, float* wy);

/* SlackProp::forward static, old interface (stereo/operators/slackprop.h:43-44,
 * slackprop.cpp:111-139): same argument order; delta = 2 as in the reference.
 * cost_volume: ih x iw x ic (ic = number of labels), output: ih x iw labels
 * written as float.  *lower_bound receives the dual bound.  Synchronises;
 * returns CCS_ERR_NAN when the cost volume holds NaN (max_iter >= 1). */
int ccs_slackprop_forward(ccs_handle_t h, const float* cost_volume, const float* wx,
                          const float* wy, float* output, int in, int ic, int ih, int iw,
                          float P1, float P2, int max_iter, float* lower_bound);

/* SlackProp::forwardVolumeOut static (stereo/operators/slackprop.h:46-47,
 * slackprop.cpp:142-173): same argument order; output (ih x iw x ic) receives
 * the reparametrised cost volume (the solver's get_C1: the final row unaries
 * of ccs_slackprop) instead of labels, delta is the caller's.  Synchronises;
 * CCS_ERR_NAN as ccs_slackprop_forward. */
int ccs_slackprop_forward_volume_out(ccs_handle_t h, const float* cost_volume, const float* wx,
                                     const float* wy, float* output, int in, int ic, int ih, int iw,
                                     float P1, float P2, int max_iter, int delta, float* lower_bound);

/* SlackProp CRF inference (the restated dual minorize-maximize over row and column
 * chains, float32 min-sum, float64 bound bookkeeping).  Call sites it replaces:
 * Stereo::computeSlackPropagation (stereo/stereo.cpp:129-162) and
 * SlackProp::forwardVolumeOut (stereo/operators/slackprop.cpp:142-173).
 * cost: H x W x D; labels: H x W int32.  unaries_out (optional, H x W x D) receives
 * the final row unaries (the reparametrised volume).  sums (optional, device,
 * 4*iters+1 doubles) receives the per-half-step chain minima S_0..S_{2 iters}
 * followed by the kept constants K_0..K_{2 iters - 1}; the bound after half-step
 * j >= 1 is K_{j-1} + S_j.  lower_bound (optional, device double) receives the
 * final bound.  Stream-ordered, no host synchronisation: ccs_slackprop_status
 * reports a NaN cost afterwards. */
int ccs_slackprop(ccs_handle_t h, const float* cost, const float* wx, const float* wy, int H,
                  int W, int D, float L1, float L2, int delta, int iters, int* labels,
                  float* unaries_out, double* sums, double* lower_bound);

/* SlackProp::forward, per-edge ("full") interface (stereo/operators/slackprop.h:40,
 * slackprop.cpp:71-108): cost_volume ih x iw x ic (ic labels), pw_volume the
 * pairwise volume 1 x pw_channels x ih x iw (NCHW planes in the reference's
 * channel order P1x, P1y, P2x, P2y; x: the edge (x,y)-(x+1,y), y: (x,y)-(x,y+1)),
 * output ih x iw labels written as float; *lower_bound the dual bound.
 * pw_channels 4 gives rho_e(d) = 0, P1_e for |d| = 1, P2_e for |d| >= 2 (delta
 * = 2); pw_channels 2 the Potts form rho_e(d) = P_e for d != 0.  The penalties
 * must satisfy 0 <= P1 <= P2 on every edge (CCS_ERR_ARG otherwise).  The
 * reference passes the volume to an un-vendored solver whose use of the levels
 * is not visible; this is the restated meaning.  Synchronises; CCS_ERR_NAN as
 * ccs_slackprop_forward. */
int ccs_slackprop_forward_pw(ccs_handle_t h, const float* cost_volume, const float* pw_volume,
                             float* output, int in, int ic, int ih, int iw, int pw_channels,
                             int max_iter, float* lower_bound);

/* The re-layout of the pairwise volume that the reference makes before its
 * solver (stereo/operators/slackprop.cpp:88-92): pen (levels x H x W x 2,
 * levels = pw_channels / 2) = pw (pw_channels x H x W) split into (2, levels)
 * with x/y fastest, levels flipped, x/y moved last:
 * pen[l][y][x][c] = pw[c + 2 (levels-1-l)][y][x].  bad (ceil(H W / 256) ints)
 * flags the pixel blocks with an edge that breaks 0 <= P1 <= P2.  Stream-ordered. */
int ccs_pairwise_penalties(ccs_handle_t h, const float* pw, int pw_channels, int H, int W, float* pen,
                           int* bad);

/* ccs_slackprop with per-edge penalties pen (levels x H x W x 2, as
 * ccs_pairwise_penalties leaves them: P1 = pen[levels-1], P2 = pen[0]; the x
 * components price the row edges, the y components the column edges) instead
 * of edge weights times L1, L2.  levels 1 or 2; delta as in ccs_slackprop.
 * Stream-ordered. */
int ccs_slackprop_pw(ccs_handle_t h, const float* cost, const float* pen, int levels, int H, int W,
                     int D, int delta, int iters, int* labels, float* unaries_out, double* sums,
                     double* lower_bound);

/* Waits for the handle's stream and returns CCS_ERR_NAN if the first row sweep
 * of the last ccs_slackprop run on this handle (or the last direction-3
 * ccs_slackprop_sweep) read a NaN cost, CCS_OK otherwise (also when no such
 * sweep ran: iters == 0).  The reference has no such check
 * (stereo/error_util.h:23-52 covers CUDA and cuDNN status codes only). */
int ccs_slackprop_status(ccs_handle_t h);

/* One half-step of ccs_slackprop on a caller-held state volume A (H x W x D)
 * that alternates between the two subproblems' unaries, so that each sweep
 * reads one volume for its chains: direction 0 sweeps the rows of the row
 * unaries in A and leaves the column unaries there (A := cost - (A -
 * minorant)); 1 sweeps the columns of those and leaves the row unaries
 * (A := (cost - A) + minorant); 2 decodes the rows of row unaries into labels;
 * 3 is the first row sweep of a run, which reads the row unaries from cost
 * (A needs no initialisation).  A run is 3, then 1, then (0, 1) per further
 * iteration, then 2.  Every sweep but 2 reads cost, which must not be A.
 * chain_sum / kept_sum (optional device doubles) receive S_j / K_j.
 * The steps of slack_prop_2D_alg::execute (stereo/stereo.cpp:153), exposed for
 * callers that schedule or time the sweeps themselves. */
int ccs_slackprop_sweep(ccs_handle_t h, const float* cost, float* A, const float* wx, const float* wy,
                        int H, int W, int D, float L1, float L2, int delta, int direction,
                        int* labels, double* chain_sum, double* kept_sum);

/* Row decode of caller-held row unaries (as a column sweep leaves them) with the
 * min-marginal gap: labels (H x W int32) = lowest argmin of the row min-marginals,
 * gap (H x W float32) = second-best minus best min-marginal, >= 0.  The gap marks
 * pixels whose label is decided by a near tie (parity checks use it); the decode
 * step of slack_prop_2D_alg (stereo/stereo.cpp:153-154) with this extra output. */
int ccs_slackprop_decode_gap(ccs_handle_t h, const float* unaries, const float* wx, int H, int W,
                             int D, float L1, float L2, int delta, int* labels, float* gap);

/* CRF energy of a labelling, accumulated in float64 (device double). */
int ccs_crf_energy(ccs_handle_t h, const float* cost, const float* wx, const float* wy, int H,
                   int W, int D, const int* labels, float L1, float L2, int delta,
                   double* energy);

/* ccs_crf_energy over the labelled part of the grid: pixels p whose mask[p]
 * (e.g. the ground truth of training) lies outside [0, D) are left out, with
 * their unary and every edge that has such an endpoint.  The energies of
 * ccs_crf_ssvm_grad.  No reference counterpart (training lived outside the repo). */
int ccs_crf_energy_masked(ccs_handle_t h, const float* cost, const float* wx, const float* wy, int H,
                          int W, int D, const int* labels, const int* mask, float L1, float L2,
                          int delta, double* energy);

/* Loss-augmented unaries for structured-SVM training: out = cost - margin [l != target]
 * per pixel (Hamming loss; targets outside [0, D) leave the pixel unchanged);
 * out may alias cost.  Running ccs_slackprop on out gives the most violating
 * labelling xhat.  No reference counterpart (training lived outside the repo). */
int ccs_crf_loss_augment(ccs_handle_t h, const float* cost, const int* target, long HW, int D, float margin,
                         float* out);

/* Structured-SVM subgradient of E(x*) - E(xhat) w.r.t. the unaries (dC, H x W x D,
 * overwritten), L1/L2 (dL, device, 2 doubles) and the edge weights (dwx, dwy,
 * H x W, optional).  Pixels whose x* lies outside [0, D) are unlabelled: both
 * energies leave out their unary and their edges (ccs_crf_energy_masked with
 * mask = x*), so dC is 0 there.  No reference counterpart (training lived
 * outside the repo). */
int ccs_crf_ssvm_grad(ccs_handle_t h, const int* xhat, const int* xstar, const float* wx,
                      const float* wy, int H, int W, int D, float L1, float L2, int delta,
                      float* dC, double* dL, float* dwx, float* dwy);

/* ---- stand-alone operator forms and layout moves ------------------------ */

/* y += bias per channel, in place.  layout CCS_LAYOUT_NCHW (planar C x H x W) or
 * CCS_LAYOUT_NHWC (channel-last with channel stride cs >= C).  Replaces
 * Bias::forward (stereo/operators/bias.cpp:58-66); the feature towers fuse it
 * into ccs_conv2d_forward instead. */
int ccs_bias_forward(ccs_handle_t h, float* y, int layout, int C, int H, int W, int cs,
                     const float* bias);

/* y = tanh(y) over count floats, in place.  Replaces Activation::forward with
 * CUDNN_ACTIVATION_TANH (stereo/operators/activation.cpp:25,33-40). */
int ccs_tanh_forward(ccs_handle_t h, float* y, long count);

/* Head of the pairwise network: out[k][p] = |sum_c w[k][c] x[p][c]|, k < K <= 4,
 * written as K planes of H x W floats (x channel-last, stride round_up(C, 8),
 * fp32 or bf16; w the K x C x 1 x 1 kernel).  Replaces the 1x1 Convolution
 * and the iu::math::abs of the pairwise net (stereo/stereonet.cpp:164,402). */
int ccs_pairwise_head(ccs_handle_t h, const void* x, int precision, int H, int W, int C, const float* w, int K,
                      float* out);

/* The two position planes of the x/y-input ColorStereoNet (out: 2 x H x W
 * floats, x plane then y plane): (x - W/2) / W and (y - H/2) / H on the padded
 * image (stereo/colorstereonet.cpp:101-112). */
int ccs_xy_planes(ccs_handle_t h, int H, int W, float* out);

/* y *= s over count floats (iu::math::mulC of the pairwise weights by lambda,
 * stereo/stereo.cpp:134-135). */
int ccs_scale_forward(ccs_handle_t h, float* y, long count, float s);

/* Softmax over the D contiguous channels of each of HW pixels, in place
 * (max-subtracted); negate != 0 also applies cost = -p.  Replaces
 * Softmax::forward (stereo/operators/softmax.cpp:30-38, CHANNEL mode, NHWC)
 * followed by cuda::negateTensor (stereo/utils.cu:289-292). */
int ccs_softmax_forward(ccs_handle_t h, float* x, long HW, int D, int negate);

/* Adjoint of cost = -softmax_D(S): dS = p (sum_k p_k g_k - g) per pixel with
 * p = -cost and g = dcost (HW x D each; dS may alias dcost).  Backward of the
 * fused epilogue of ccs_correlation_volume; no reference counterpart (the
 * reference has no softmax backward). */
int ccs_neg_softmax_backward(ccs_handle_t h, const float* cost, const float* dcost, long HW, int D, float* dS);

/* RGBA float image (16-byte aligned, row pitch in floats) -> planar C x H x W
 * (C <= 4 leading channels).  The copy at the top of ColorStereoNet::predict
 * (stereo/colorstereonet.cpp:53-54). */
int ccs_c4_to_planar(ccs_handle_t h, const float* c4, int H, int W, long pitch, int C, float* out);

/* Channel-last feature map (stride cs, CCS_FP32 or CCS_BF16) -> planar fp32
 * C x H x W, and planar fp32 -> channel-last fp32 (pad channels zeroed).  Layout
 * moves for callers that hold NCHW tensors (iu::TensorGpu_32f NCHW). */
int ccs_nhwc_to_nchw(ccs_handle_t h, const void* in, int precision, int C, int cs, int H, int W,
                     float* out);
int ccs_nchw_to_nhwc(ccs_handle_t h, const float* in, int C, int cs, int H, int W, float* out);

/* ---- decode -------------------------------------------------------------- */

/* First strict minimum over D per pixel -> disp_min + index * disp_step
 * (stereo/stereo.cu:348-370, 975-994).  out: H x W float32. */
int ccs_argmin(ccs_handle_t h, const float* cost, int H, int W, int D, float disp_min,
               float disp_step, float* out);

/* QuadDirect sub-pixel refinement, in place on the H x W disparity map
 * (Stereo::fuseQuadFitDirect, stereo/stereo.cpp:406-409 and
 * stereo/stereo.cu:924-942,1172-1195): z = (int)((d - disp_min) / disp_step);
 * for 0 < z < D - 1, d += g / k with g = (c[z-1] - c[z+1]) / (2 step) and
 * k = max(1e-3, (c[z-1] + c[z+1] - 2 c[z]) / step^2), c the pixel's D volume
 * values (H x W x D); other pixels keep d. */
int ccs_quad_fit_direct(ccs_handle_t h, float* disparity, const float* volume, int H, int W, int D,
                        float disp_min, float disp_step);

/* Left-right check (Stereo::performLeftRightCheck, stereo/stereo.cpp:448-455 and
 * stereo/stereo.cu:803-817,1198-1225): out = left where |left + R(x - left)| <= th,
 * else -1; right holds the right-to-left disparities (negative), sampled with
 * linear interpolation along the row, clamped at the borders. */
int ccs_left_right_check(ccs_handle_t h, const float* left, const float* right, int H, int W, float th,
                         float* out);

/* label * disp_step + disp_min (stereo/stereo.cu:841-850, 1229-1246). */
int ccs_remap(ccs_handle_t h, const int* labels, int H, int W, float disp_min, float disp_step,
              float* out);

#ifdef __cplusplus
}
#endif
#endif
