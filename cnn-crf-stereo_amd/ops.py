"""Torch-tensor front end of the C ABI: each function allocates its outputs with
torch (device memory only) and runs the HIP kernels of libccs.so on the
handle's stream.  Feature maps are H x W x Cs tensors (Cs = C rounded up to 8),
float32 or bfloat16; cost volumes are H x W x D float32."""
import ctypes

import numpy as np
import torch

from . import capi
from .capi import call, ptr

PRECISIONS = {"fp32": capi.FP32, "bf16": capi.BF16}


def _dtype(precision):
    return torch.bfloat16 if precision == "bf16" else torch.float32


def _cuda(a, dtype=torch.float32):
    if isinstance(a, torch.Tensor):
        return a.to(device="cuda", dtype=dtype).contiguous()
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device="cuda")


def pad_normalize(h, img, pad, stats=None):
    """img: C x H x W float32 (cuda) -> C x (H+2p) x (W+2p) normalised, mirror padded."""
    C, H, W = img.shape
    out = torch.empty((C, H + 2 * pad, W + 2 * pad), dtype=torch.float32, device="cuda")
    call("ccs_pad_normalize", h.h, ptr(img), C, H, W, pad, ptr(out), ptr(stats))
    return out


def rgb_to_gray(h, rgb):
    _, H, W = rgb.shape
    out = torch.empty((H, W), dtype=torch.float32, device="cuda")
    call("ccs_rgb_to_gray", h.h, ptr(rgb), H, W, ptr(out))
    return out


class ConvLayer:
    """One tower layer: weights (K, C, KH, KW) float32 + bias, optionally packed for MFMA."""

    def __init__(self, h, w, b, precision="fp32", act=True):
        self.w = _cuda(w)
        self.b = _cuda(b) if b is not None else None
        self.K, self.C, self.KH, self.KW = self.w.shape
        self.precision = precision
        self.act = int(act)
        self.packed = None
        cs = capi.feature_stride(self.C)
        if self._mfma_shape(cs):
            nbytes = capi.lib().ccs_conv_packed_bytes(self.K, self.C, self.KH, self.KW, PRECISIONS[precision])
            self.packed = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
            call("ccs_conv_pack_weights", h.h, ptr(self.w), self.K, self.C, self.KH, self.KW,
                 PRECISIONS[precision], ptr(self.packed))

    def _mfma_shape(self, cs):
        # Mirrors mfma_supported() in csrc/conv.hip: output stride must equal the input stride.
        nt = (self.K + 15) // 16
        return ((self.KH, self.KW) in ((3, 3), (2, 2)) and capi.feature_stride(self.K) == cs
                and ((cs == 104 and nt == 7) or (cs == 8 and nt == 1)))

    def __call__(self, h, x, layout="nhwc", out=None):
        if layout == "nchw":
            _, H, W = x.shape
        else:
            H, W, _ = x.shape
        Ho, Wo = H - self.KH + 1, W - self.KW + 1
        if out is None:
            out = torch.empty((Ho, Wo, capi.feature_stride(self.K)), dtype=_dtype(self.precision), device="cuda")
        lay = capi.LAYOUT_NCHW if layout == "nchw" else capi.LAYOUT_NHWC
        if self.packed is not None and layout == "nhwc":
            call("ccs_conv_forward_packed", h.h, ptr(x), lay, H, W, self.C, ptr(self.packed), ptr(self.b),
                 self.K, self.KH, self.KW, self.act, PRECISIONS[self.precision], ptr(out))
        else:
            call("ccs_conv2d_forward", h.h, ptr(x), lay, H, W, self.C, ptr(self.w), ptr(self.b),
                 self.K, self.KH, self.KW, self.act, PRECISIONS[self.precision], ptr(out))
        return out


class Tower:
    """Siamese feature tower (shared weights for both images)."""

    def __init__(self, h, params, precision="fp32"):
        self.layers = [ConvLayer(h, w, b, precision) for (w, b) in params]
        self.precision = precision
        self.pad = sum(l.KH - 1 for l in self.layers) // 2

    def __call__(self, h, img_padded):
        x = self.layers[0](h, img_padded, layout="nchw")
        for layer in self.layers[1:]:
            x = layer(h, x)
        return x


def conv2d_backward(h, x, w, b, dy, act=True):
    """float32 feature maps; returns dx (H x W x Cs), dw, db."""
    H, W, _ = x.shape
    K, C, KH, KW = w.shape
    dx = torch.empty((H, W, capi.feature_stride(C)), dtype=torch.float32, device="cuda")
    dw = torch.empty_like(w)
    db = torch.empty((K,), dtype=torch.float32, device="cuda")
    call("ccs_conv2d_backward", h.h, ptr(x), H, W, C, ptr(w), ptr(b), K, KH, KW, int(act), ptr(dy),
         ptr(dx), ptr(dw), ptr(db))
    return dx, dw, db


def conv2d_backward_stored(h, x, y, w, dy, act=True, want_dx=True):
    """Adjoint from the layer's input x and activated output y (no recompute).
    x: H x W x Cs (float32, or y's dtype), y and dy: Ho x Wo x Ks (float32 or
    bfloat16); returns dx (x's shape, y's dtype, or None), dw, db (float32)."""
    H, W, _ = x.shape
    K, C, KH, KW = w.shape
    xp = capi.BF16 if x.dtype == torch.bfloat16 else capi.FP32
    yp = capi.BF16 if y.dtype == torch.bfloat16 else capi.FP32
    dx = torch.empty((H, W, capi.feature_stride(C)), dtype=y.dtype, device="cuda") if want_dx else None
    dw = torch.empty((K, C, KH, KW), dtype=torch.float32, device="cuda")
    db = torch.empty((K,), dtype=torch.float32, device="cuda")
    call("ccs_conv2d_backward_stored", h.h, ptr(x), xp, ptr(y), yp, H, W, C, ptr(w), K, KH, KW, int(act), ptr(dy),
         ptr(dx), ptr(dw), ptr(db))
    return dx, dw, db


def correlation_volume(h, f0, f1, C, disp_min, disp_step, D, epilogue="neg_softmax", out=None):
    H, W, _ = f0.shape
    precision = "bf16" if f0.dtype == torch.bfloat16 else "fp32"
    if out is None:
        out = torch.empty((H, W, D), dtype=torch.float32, device="cuda")
    ep = capi.EPILOGUE_NEG_SOFTMAX if epilogue == "neg_softmax" else capi.EPILOGUE_RAW
    call("ccs_correlation_volume", h.h, ptr(f0), ptr(f1), PRECISIONS[precision], H, W, C, int(disp_min),
         int(disp_step), D, ep, ptr(out))
    return out


def correlation_volume_backward(h, f0, f1, C, disp_min, disp_step, D, G):
    """df0, df1 in the feature maps' dtype (float32 or bfloat16)."""
    H, W, Cs = f0.shape
    precision = "bf16" if f0.dtype == torch.bfloat16 else "fp32"
    df0 = torch.empty_like(f0)
    df1 = torch.empty_like(f1)
    call("ccs_correlation_volume_backward", h.h, ptr(f0), ptr(f1), PRECISIONS[precision], H, W, C, int(disp_min),
         int(disp_step), D, ptr(G), ptr(df0), ptr(df1))
    return df0, df1


def correlation_forward_nchw(h, f0, f1, disparities, layout="nhwc", rect_corr=0):
    """The reference's static StereoCorrelation::forward on N x C x H x W inputs."""
    N, C, H, W = f0.shape
    D = disparities.numel()
    if layout == "nhwc":
        out = torch.empty((N, H, W, D), dtype=torch.float32, device="cuda")
    else:
        out = torch.empty((N, D, H, W), dtype=torch.float32, device="cuda")
    call("ccs_correlation_forward", h.h, ptr(f0), ptr(f1), ptr(disparities), ptr(out), N, C, H, W, D,
         capi.LAYOUT_NHWC if layout == "nhwc" else capi.LAYOUT_NCHW, rect_corr)
    return out


def correlation_backward_nchw(h, G, f0, f1, disparities, layout="nhwc"):
    N, C, H, W = f0.shape
    D = disparities.numel()
    d0 = torch.empty_like(f0)
    d1 = torch.empty_like(f1)
    call("ccs_correlation_backward", h.h, ptr(d0), ptr(d1), ptr(G), ptr(f0), ptr(f1), ptr(disparities), N, C, H, W,
         D, capi.LAYOUT_NHWC if layout == "nhwc" else capi.LAYOUT_NCHW)
    return d0, d1


def softmax_xent(h, S, target):
    H, W, D = S.shape
    loss = torch.zeros(1, dtype=torch.float64, device="cuda")
    dS = torch.empty_like(S)
    call("ccs_softmax_xent", h.h, ptr(S), H * W, D, ptr(target), ptr(loss), ptr(dS))
    return loss, dS


def edge_weights(h, gray, sigma=0.7, alpha=10.0, beta=1.0, lam=1.0):
    H, W = gray.shape
    wx = torch.empty((H, W), dtype=torch.float32, device="cuda")
    wy = torch.empty((H, W), dtype=torch.float32, device="cuda")
    call("ccs_edge_weights", h.h, ptr(gray), H, W, float(sigma), float(alpha), float(beta), float(lam),
         ptr(wx), ptr(wy))
    return wx, wy


def slackprop(h, cost, wx, wy, L1, L2, delta, iters, want_unaries=False):
    """Returns dict(labels int32 HxW, sums float64 (4 iters + 1), lb float64 (1,), [unaries])."""
    H, W, D = cost.shape
    labels = torch.empty((H, W), dtype=torch.int32, device="cuda")
    sums = torch.zeros(4 * iters + 1, dtype=torch.float64, device="cuda")
    lb = torch.zeros(1, dtype=torch.float64, device="cuda")
    un = torch.empty_like(cost) if want_unaries else None
    call("ccs_slackprop", h.h, ptr(cost), ptr(wx), ptr(wy), H, W, D, float(L1), float(L2), int(delta), int(iters),
         ptr(labels), ptr(un), ptr(sums), ptr(lb))
    out = dict(labels=labels, sums=sums, lb=lb)
    if want_unaries:
        out["unaries"] = un
    return out


def slackprop_sweep(h, cost, A, wx, wy, L1, L2, delta, direction, labels=None, chain_sum=None, kept_sum=None):
    """One half-step (0 rows, 1 columns, 2 decode, 3 the first row sweep of a run,
    which reads the unaries from cost so that A needs no copy of it) on the
    caller's state volume A: a row sweep reads the row unaries there and leaves
    the column unaries cost - A', a column sweep the reverse; the decode reads
    row unaries (as a column sweep leaves them)."""
    H, W, D = cost.shape
    call("ccs_slackprop_sweep", h.h, ptr(cost), ptr(A), ptr(wx), ptr(wy), H, W, D, float(L1), float(L2),
         int(delta), int(direction), ptr(labels), ptr(chain_sum), ptr(kept_sum))


def slackprop_decode_gap(h, unaries, wx, L1, L2, delta):
    H, W, D = unaries.shape
    labels = torch.empty((H, W), dtype=torch.int32, device="cuda")
    gap = torch.empty((H, W), dtype=torch.float32, device="cuda")
    call("ccs_slackprop_decode_gap", h.h, ptr(unaries), ptr(wx), H, W, D, float(L1), float(L2), int(delta),
         ptr(labels), ptr(gap))
    return labels, gap


def slackprop_forward_static(h, cost, wx, wy, P1, P2, max_iter=5):
    """The reference's SlackProp::forward (old interface, delta = 2): labels as float and the bound."""
    H, W, D = cost.shape
    out = torch.empty((H, W), dtype=torch.float32, device="cuda")
    lb = ctypes.c_float()
    call("ccs_slackprop_forward", h.h, ptr(cost), ptr(wx), ptr(wy), ptr(out), 1, D, H, W, float(P1), float(P2),
         int(max_iter), ctypes.byref(lb))
    return out, lb.value


def slackprop_forward_volume_out(h, cost, wx, wy, P1, P2, max_iter, delta):
    """The reference's SlackProp::forwardVolumeOut: the reparametrised volume and the bound."""
    H, W, D = cost.shape
    out = torch.empty_like(cost)
    lb = ctypes.c_float()
    call("ccs_slackprop_forward_volume_out", h.h, ptr(cost), ptr(wx), ptr(wy), ptr(out), 1, D, H, W, float(P1),
         float(P2), int(max_iter), int(delta), ctypes.byref(lb))
    return out, lb.value


def pairwise_penalties(h, pw):
    """The reference's re-layout of a pairwise volume (2 levels x H x W) into the
    per-edge penalties (levels x H x W x 2) and the per-block validity flags."""
    C2, H, W = pw.shape
    pen = torch.empty((C2 // 2, H, W, 2), dtype=torch.float32, device="cuda")
    bad = torch.empty(((H * W + 255) // 256,), dtype=torch.int32, device="cuda")
    call("ccs_pairwise_penalties", h.h, ptr(pw), C2, H, W, ptr(pen), ptr(bad))
    return pen, bad


def slackprop_pw(h, cost, pen, delta, iters, want_unaries=False):
    """ccs_slackprop with per-edge penalties pen (levels x H x W x 2)."""
    H, W, D = cost.shape
    labels = torch.empty((H, W), dtype=torch.int32, device="cuda")
    sums = torch.zeros(4 * iters + 1, dtype=torch.float64, device="cuda")
    lb = torch.zeros(1, dtype=torch.float64, device="cuda")
    un = torch.empty_like(cost) if want_unaries else None
    call("ccs_slackprop_pw", h.h, ptr(cost), ptr(pen), pen.shape[0], H, W, D, int(delta), int(iters), ptr(labels),
         ptr(un), ptr(sums), ptr(lb))
    out = dict(labels=labels, sums=sums, lb=lb)
    if want_unaries:
        out["unaries"] = un
    return out


def slackprop_forward_pw(h, cost, pw, max_iter=5):
    """The reference's SlackProp::forward(costVolume, pwVolume, output, max_iter): labels as float and the bound."""
    H, W, D = cost.shape
    out = torch.empty((H, W), dtype=torch.float32, device="cuda")
    lb = ctypes.c_float()
    call("ccs_slackprop_forward_pw", h.h, ptr(cost), ptr(pw), ptr(out), 1, D, H, W, pw.shape[0], int(max_iter),
         ctypes.byref(lb))
    return out, lb.value


def slackprop_status(h):
    """Raises CCSError (status ERR_NAN) if the last CRF run on h read a NaN cost."""
    call("ccs_slackprop_status", h.h)


def crf_energy(h, cost, wx, wy, labels, L1, L2, delta, mask=None):
    """CRF energy of labels (float64 on the device); with mask, only over the
    pixels whose mask label lies in [0, D)."""
    H, W, D = cost.shape
    e = torch.zeros(1, dtype=torch.float64, device="cuda")
    if mask is None:
        call("ccs_crf_energy", h.h, ptr(cost), ptr(wx), ptr(wy), H, W, D, ptr(labels), float(L1), float(L2),
             int(delta), ptr(e))
    else:
        call("ccs_crf_energy_masked", h.h, ptr(cost), ptr(wx), ptr(wy), H, W, D, ptr(labels), ptr(mask),
             float(L1), float(L2), int(delta), ptr(e))
    return e


def crf_ssvm_grad(h, xhat, xstar, wx, wy, D, L1, L2, delta):
    H, W = xhat.shape
    dC = torch.empty((H, W, D), dtype=torch.float32, device="cuda")
    dL = torch.zeros(2, dtype=torch.float64, device="cuda")
    dwx = torch.empty((H, W), dtype=torch.float32, device="cuda")
    dwy = torch.empty((H, W), dtype=torch.float32, device="cuda")
    call("ccs_crf_ssvm_grad", h.h, ptr(xhat), ptr(xstar), ptr(wx), ptr(wy), H, W, D, float(L1), float(L2),
         int(delta), ptr(dC), ptr(dL), ptr(dwx), ptr(dwy))
    return dC, dL, dwx, dwy


def argmin(h, cost, disp_min=0.0, disp_step=1.0):
    H, W, D = cost.shape
    out = torch.empty((H, W), dtype=torch.float32, device="cuda")
    call("ccs_argmin", h.h, ptr(cost), H, W, D, float(disp_min), float(disp_step), ptr(out))
    return out


def remap(h, labels, disp_min=0.0, disp_step=1.0):
    H, W = labels.shape
    out = torch.empty((H, W), dtype=torch.float32, device="cuda")
    call("ccs_remap", h.h, ptr(labels), H, W, float(disp_min), float(disp_step), ptr(out))
    return out
