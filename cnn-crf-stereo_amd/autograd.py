"""torch.autograd.Function wrappers over the C-ABI adjoints, and a training step.

Every forward and backward computation runs in libccs.so; torch supplies device
memory, the stream and gradient accumulation across the two towers (the
weights are shared) and across the frames of a batch.

  Tower         one feature tower, checkpointed: ccs_conv2d_forward per layer
                (float32 or bfloat16 maps, float32 master weights), saving only
                the padded image; the backward recomputes the layer outputs and
                runs ccs_conv2d_backward_stored from the last layer down,
                releasing each output after its adjoint
  ConvTanh      one layer, store-everything: ccs_conv2d_forward /
                ccs_conv2d_backward (float32; keeps the layer input, recomputes
                the output in the backward) -- the plain form the checkpointed
                tower is compared against
  Correlation   ccs_correlation_volume (raw scores) / ccs_correlation_volume_backward
                (each pixel gathers its disparity band: no atomics)
  NegSoftmax    ccs_softmax_forward(negate) / ccs_neg_softmax_backward
  SoftmaxXent   ccs_softmax_xent: loss and dS in one pass
  SsvmHinge     ccs_crf_loss_augment + ccs_slackprop (most violating labelling),
                ccs_crf_energy, ccs_crf_ssvm_grad (subgradient w.r.t. unaries,
                L1/L2 and the edge weights)

The reference has no training code of its own (its weights were trained
elsewhere); these adjoints have no reference counterpart.
"""
import torch

from . import ops
from .capi import BF16, EPILOGUE_RAW, FP32, LAYOUT_NCHW, LAYOUT_NHWC, call, feature_stride, ptr

_DTYPE = {"fp32": torch.float32, "bf16": torch.bfloat16}
_CODE = {"fp32": FP32, "bf16": BF16}


def _zeros_hwc(H, W, C, dtype=torch.float32):
    return torch.zeros((H, W, feature_stride(C)), dtype=dtype, device="cuda")


def _tower_forward(h, img, ws, bs, precision, keep):
    """Layer outputs of one tower from the padded planar image: all of them
    (keep) or the last one only (each earlier map is released when the next
    layer has been queued)."""
    x, outs = img, []
    for l, (w, b) in enumerate(zip(ws, bs)):
        K, C, KH, KW = w.shape
        H, W = x.shape[1:] if l == 0 else x.shape[:2]
        y = torch.empty((H - KH + 1, W - KW + 1, feature_stride(K)), dtype=_DTYPE[precision], device="cuda")
        call("ccs_conv2d_forward", h.h, ptr(x), LAYOUT_NCHW if l == 0 else LAYOUT_NHWC, H, W, C, ptr(w), ptr(b), K,
             KH, KW, 1, _CODE[precision], ptr(y))
        if keep:
            outs.append(y)
        x = y
    return outs if keep else [x]


class Tower(torch.autograd.Function):
    """Feature tower y = layer_L(... layer_1(img)), fused conv + bias + tanh per
    layer, in float32 or bfloat16 (MFMA, float32 accumulation) from float32
    master weights.  Checkpointed: only the padded image (C x Hp x Wp) is
    saved.  The backward recomputes the layer outputs in the forward's
    precision (the same kernels, so the same values), then runs the adjoints
    from the last layer down; tanh' comes from the kept outputs, and each
    output is released once its layer's adjoint is queued.  Peak memory is one
    tower's maps during the backward instead of every layer of both towers
    from the forward to the backward."""

    @staticmethod
    def forward(ctx, img, precision, h, *wb):
        ws, bs = wb[0::2], wb[1::2]
        y = _tower_forward(h, img, ws, bs, precision, keep=False)[-1]
        ctx.save_for_backward(img, *wb)
        ctx.precision, ctx.h = precision, h
        return y

    @staticmethod
    def backward(ctx, dy):
        img, *wb = ctx.saved_tensors
        ws, bs = wb[0::2], wb[1::2]
        h, precision = ctx.h, ctx.precision
        acts = _tower_forward(h, img, ws, bs, precision, keep=True)
        C0, Hp, Wp = img.shape
        x0 = _zeros_hwc(Hp, Wp, C0)  # the first layer's adjoint reads the image channel-last
        call("ccs_nchw_to_nhwc", h.h, ptr(img.contiguous()), C0, feature_stride(C0), Hp, Wp, ptr(x0))
        grads = [None] * len(wb)
        g = dy.to(_DTYPE[precision]).contiguous()
        for l in reversed(range(len(ws))):
            x = acts[l - 1] if l > 0 else x0
            dx, dw, db = ops.conv2d_backward_stored(h, x, acts[l], ws[l], g, True, want_dx=l > 0)
            grads[2 * l], grads[2 * l + 1] = dw, db
            acts[l] = None  # this layer's output is no longer needed
            g = dx
        return (None, None, None, *grads)


class ConvTanh(torch.autograd.Function):
    """y = tanh(conv(x, w) + b), true convolution, channel-last fp32 maps.
    x: planar 1 x H x W image (first layer) or H x W x stride(C) features."""

    @staticmethod
    def forward(ctx, x, w, b, h, planar=False):
        K, C, KH, KW = w.shape
        if planar:
            _, H, W = x.shape
        else:
            H, W = x.shape[:2]
        y = _zeros_hwc(H - KH + 1, W - KW + 1, K)
        w = w.contiguous()
        b = b.contiguous()
        call("ccs_conv2d_forward", h.h, ptr(x), LAYOUT_NCHW if planar else LAYOUT_NHWC, H, W, C, ptr(w), ptr(b),
             K, KH, KW, 1, FP32, ptr(y))
        if planar:  # the backward reads channel-last input
            xh = _zeros_hwc(H, W, C)
            call("ccs_nchw_to_nhwc", h.h, ptr(x.contiguous()), C, feature_stride(C), H, W, ptr(xh))
        else:
            xh = x
        ctx.save_for_backward(xh, w, b)
        ctx.h, ctx.planar = h, planar
        return y

    @staticmethod
    def backward(ctx, dy):
        xh, w, b = ctx.saved_tensors
        h = ctx.h
        K, C, KH, KW = w.shape
        H, W = xh.shape[:2]
        dy = dy.contiguous()
        if dy.shape[2] > K:
            dy = dy.clone()
            dy[:, :, K:] = 0
        dw = torch.empty_like(w)
        db = torch.empty_like(b)
        need_dx = ctx.needs_input_grad[0] and not ctx.planar
        dx = _zeros_hwc(H, W, C) if need_dx else None
        call("ccs_conv2d_backward", h.h, ptr(xh), H, W, C, ptr(w), ptr(b), K, KH, KW, 1, ptr(dy), ptr(dx), ptr(dw),
             ptr(db))
        return dx, dw, db, None, None


class Correlation(torch.autograd.Function):
    """S[h][w][k] = <f0[h][w], f1[h][w - (disp_min + k disp_step)]> (raw scores);
    float32 or bfloat16 maps (MFMA), float32 scores."""

    @staticmethod
    def forward(ctx, f0, f1, C, disp_min, disp_step, D, h):
        S = ops.correlation_volume(h, f0, f1, C, disp_min, disp_step, D, epilogue="raw")
        ctx.save_for_backward(f0, f1)
        ctx.args = (C, disp_min, disp_step, D, h)
        return S

    @staticmethod
    def backward(ctx, G):
        f0, f1 = ctx.saved_tensors
        C, disp_min, disp_step, D, h = ctx.args
        df0, df1 = ops.correlation_volume_backward(h, f0, f1, C, disp_min, disp_step, D, G.contiguous())
        return df0, df1, None, None, None, None, None


class NegSoftmax(torch.autograd.Function):
    """cost = -softmax_D(S)."""

    @staticmethod
    def forward(ctx, S, h):
        H, W, D = S.shape
        cost = S.detach().clone()
        call("ccs_softmax_forward", h.h, ptr(cost), H * W, D, 1)
        ctx.save_for_backward(cost)
        ctx.h = h
        return cost

    @staticmethod
    def backward(ctx, g):
        (cost,) = ctx.saved_tensors
        H, W, D = cost.shape
        dS = torch.empty_like(cost)
        call("ccs_neg_softmax_backward", ctx.h.h, ptr(cost), ptr(g.contiguous()), H * W, D, ptr(dS))
        return dS, None


class SoftmaxXent(torch.autograd.Function):
    """sum over pixels of -log softmax_D(S)[target]; targets outside [0, D) are ignored."""

    @staticmethod
    def forward(ctx, S, target, h):
        H, W, D = S.shape
        loss = torch.zeros(1, dtype=torch.float64, device="cuda")
        dS = torch.empty_like(S)
        call("ccs_softmax_xent", h.h, ptr(S.contiguous()), H * W, D, ptr(target), ptr(loss), ptr(dS))
        ctx.save_for_backward(dS)
        return loss.float().reshape(())

    @staticmethod
    def backward(ctx, g):
        (dS,) = ctx.saved_tensors
        return dS * g, None, None


class SsvmHinge(torch.autograd.Function):
    """Structured-SVM hinge  E(x*) - [E_aug(xhat)]  with xhat the SlackProp
    minimum of the loss-augmented energy; params = (L1, L2).  Pixels without
    ground truth (target outside [0, D)) drop out of both energies with their
    unaries and their edges (the masked energies of ccs_crf_ssvm_grad)."""

    @staticmethod
    def forward(ctx, cost, wx, wy, params, target, iters, delta, margin, h):
        H, W, D = cost.shape
        L1, L2 = float(params[0]), float(params[1])
        aug = torch.empty_like(cost)
        call("ccs_crf_loss_augment", h.h, ptr(cost), ptr(target), H * W, D, float(margin), ptr(aug))
        r = ops.slackprop(h, aug, wx, wy, L1, L2, delta, iters)
        xhat = r["labels"]
        xstar = target.clamp(0, D - 1)  # (only read where target is valid: the mask drops the rest)
        e_star = ops.crf_energy(h, cost, wx, wy, xstar, L1, L2, delta, mask=target)
        e_hat = ops.crf_energy(h, aug, wx, wy, xhat, L1, L2, delta, mask=target)
        ctx.save_for_backward(xhat, target, wx, wy)
        ctx.args = (D, L1, L2, delta, h)
        return (e_star - e_hat).float().reshape(())

    @staticmethod
    def backward(ctx, g):
        xhat, target, wx, wy = ctx.saved_tensors
        D, L1, L2, delta, h = ctx.args
        dC, dL, dwx, dwy = ops.crf_ssvm_grad(h, xhat, target, wx, wy, D, L1, L2, delta)
        return dC * g, dwx * g, dwy * g, (dL.float() * g), None, None, None, None, None


class StereoModel:
    """Feature tower weights (shared by both eyes, float32 masters) and the
    learned L1, L2, as leaf tensors with requires_grad.

    precision "fp32" or "bf16" is the towers' and the correlation's arithmetic;
    checkpoint=True runs each tower as one checkpointed Tower, False as
    store-everything ConvTanh layers (float32 only)."""

    def __init__(self, params, L1=0.5, L2=2.0, precision="fp32", checkpoint=True):
        if precision not in _DTYPE or (precision == "bf16" and not checkpoint):
            raise ValueError("precision must be fp32, or bf16 with checkpoint=True")
        self.weights = [(torch.as_tensor(w, device="cuda").clone().requires_grad_(),
                         torch.as_tensor(b, device="cuda").clone().requires_grad_()) for (w, b) in params]
        self.crf = torch.tensor([L1, L2], dtype=torch.float32, device="cuda", requires_grad=True)
        self.pad = sum(w.shape[2] - 1 for (w, _) in self.weights) // 2
        self.precision, self.checkpoint = precision, checkpoint

    def parameters(self):
        return [t for wb in self.weights for t in wb] + [self.crf]

    def tower(self, h, img_padded):
        if self.checkpoint:
            return Tower.apply(img_padded, self.precision, h, *[t for wb in self.weights for t in wb])
        x = img_padded
        for i, (w, b) in enumerate(self.weights):
            x = ConvTanh.apply(x, w, b, h, i == 0)
        return x

    def scores(self, h, left, right, D, disp_min=0, disp_step=1):
        """Raw correlation scores S (H x W x D) of one frame (images 1 x H x W)."""
        f = [self.tower(h, ops.pad_normalize(h, img, self.pad)) for img in (left, right)]
        C = self.weights[-1][0].shape[0]
        return Correlation.apply(f[0], f[1], C, disp_min, disp_step, D, h)


def frame_loss(model, h, left, right, target, D, iters=4, delta=2, hinge_weight=1.0, sigma=0.7, alpha=10.0,
               beta=1.0, lam=1.0):
    """Unary cross entropy of softmax(S) at the true labels plus the CRF hinge
    w.r.t. the pairwise parameters (the unaries enter the hinge detached), both
    per pixel."""
    H, W = left.shape[1:]
    S = model.scores(h, left, right, D)
    n = max(1, int((target >= 0).sum().item()))
    loss = SoftmaxXent.apply(S, target, h) / n
    if hinge_weight:
        cost = NegSoftmax.apply(S, h).detach()
        wx, wy = ops.edge_weights(h, left[0], sigma, alpha, beta, lam)
        loss = loss + hinge_weight * SsvmHinge.apply(cost, wx, wy, model.crf, target, iters, delta, 1.0, h) / (H * W)
    return loss


def gradients(model, h, batch, D, **kw):
    """Mean loss over a batch of (left, right, target) frames and the gradients
    of every parameter (float32 copies), without updating the model."""
    for p in model.parameters():
        p.grad = None
    total = 0.0
    for (left, right, target) in batch:
        loss = frame_loss(model, h, left, right, target, D, **kw) / len(batch)
        loss.backward()
        total += float(loss.detach())
    grads = [p.grad.detach().clone() if p.grad is not None else torch.zeros_like(p) for p in model.parameters()]
    for p in model.parameters():
        p.grad = None
    return total, grads


def training_step(model, h, batch, D, lr=1e-3, **kw):
    """One SGD step over a batch of (left, right, target) frames; returns the mean loss."""
    for p in model.parameters():
        p.grad = None
    total = 0.0
    for (left, right, target) in batch:
        loss = frame_loss(model, h, left, right, target, D, **kw) / len(batch)
        loss.backward()
        total += float(loss.detach())
    with torch.no_grad():
        for p in model.parameters():
            if p.grad is not None:
                p -= lr * p.grad
        model.crf.clamp_(min=0.0)
        model.crf[1] = torch.maximum(model.crf[1], model.crf[0])  # keep 0 <= L1 <= L2
    return total
