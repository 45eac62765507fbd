"""GPU parity: every HIP kernel through the C ABI against the FP64 oracle on the
same seeded inputs.  Tolerances are stated per test:
  - float32 kernels: absolute error <= 2e-4 on O(1) values (float32 accumulation);
  - bfloat16 MFMA kernels: the oracle gets the same bf16-rounded inputs and
    weights, so only accumulation order and the bf16 output rounding differ
    (absolute <= 8e-3 on tanh outputs, relative 1e-2 on correlations);
  - integer/index outputs (argmin, remap, labels) are compared exactly, except
    CRF labels, which may differ only where the oracle's min-marginal gap is
    below 1e-3;
  - CRF energies and bounds: relative 1e-5.
"""
import numpy as np
import pytest
import torch

import oracle as O
from cnn_crf_stereo_amd import capi, ops, synth
from cnn_crf_stereo_amd.capi import call, ptr

pytestmark = pytest.mark.gpu


def bf16_round(a):
    return torch.as_tensor(np.asarray(a, dtype=np.float32)).to(torch.bfloat16).to(torch.float32).numpy()


def to_hwc(x_chw, cs, dtype=torch.float32):
    C, H, W = x_chw.shape
    out = torch.zeros((H, W, cs), dtype=dtype, device="cuda")
    out[:, :, :C] = torch.as_tensor(np.ascontiguousarray(x_chw.transpose(1, 2, 0)), dtype=torch.float32).to(dtype)
    return out


def from_hwc(t, C):
    return t[:, :, :C].float().cpu().numpy().transpose(2, 0, 1).astype(np.float64)


def glorot(rng, K, C, KH, KW):
    a = np.sqrt(6.0 / (C * KH * KW + K * KH * KW))
    return rng.uniform(-a, a, size=(K, C, KH, KW)).astype(np.float32)


# ---------------------------------------------------------------- input prep

def test_pad_normalize(handle):
    rng = np.random.default_rng(1)
    img = rng.uniform(0, 1, size=(3, 21, 33)).astype(np.float32)
    p = 4
    g = ops.pad_normalize(handle, torch.as_tensor(img, device="cuda"), p)
    ref, mean, std = O.normalize(O.pad_symmetric(img, p))
    np.testing.assert_allclose(g.cpu().numpy(), ref, atol=2e-5, rtol=0)


def test_rgb_to_gray(handle):
    rng = np.random.default_rng(2)
    rgb = rng.uniform(0, 1, size=(3, 17, 29)).astype(np.float32)
    g = ops.rgb_to_gray(handle, torch.as_tensor(rgb, device="cuda"))
    np.testing.assert_allclose(g.cpu().numpy(), O.rgb2gray(rgb), atol=1e-6)


def test_edge_weights(handle):
    rng = np.random.default_rng(3)
    gray = synth.gaussian_blur(rng.uniform(0, 1, size=(1, 37, 53)), 1.0)[0].astype(np.float32)
    for (alpha, beta, lam) in [(10.0, 1.0, 1.0), (20.0, 0.8, 0.4), (0.0, 1.0, 0.5)]:
        wx, wy = ops.edge_weights(handle, torch.as_tensor(gray, device="cuda"), 0.7, alpha, beta, lam)
        rx, ry = O.edge_weights(gray, 0.7, alpha, beta, lam)
        np.testing.assert_allclose(wx.cpu().numpy(), rx, rtol=2e-5, atol=1e-6)
        np.testing.assert_allclose(wy.cpu().numpy(), ry, rtol=2e-5, atol=1e-6)


# ---------------------------------------------------------------- convolution

@pytest.mark.parametrize("precision", ["fp32", "bf16"])
@pytest.mark.parametrize("C,K,k", [(100, 100, 3), (100, 100, 2), (64, 64, 3), (8, 8, 3), (8, 8, 2), (5, 12, 3)])
def test_conv_layer(handle, precision, C, K, k):
    rng = np.random.default_rng(10 + C + K + k)
    H, W = 9, 70 + k
    x = rng.uniform(-1, 1, size=(C, H, W)).astype(np.float32)
    w = glorot(rng, K, C, k, k)
    b = rng.uniform(-0.1, 0.1, size=K).astype(np.float32)
    dt = torch.bfloat16 if precision == "bf16" else torch.float32
    layer = ops.ConvLayer(handle, w, b, precision)
    y = layer(handle, to_hwc(x, capi.feature_stride(C), dt))
    if precision == "bf16":
        x, w = bf16_round(x), bf16_round(w)
        tol = 8e-3
    else:
        tol = 2e-4
    ref = O.conv2d(x, w, b, True)
    got = from_hwc(y, K)
    np.testing.assert_allclose(got, ref, atol=tol, rtol=0)
    pad = y[:, :, K:].float()
    assert torch.count_nonzero(pad).item() == 0, "padding channels must stay zero"


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
@pytest.mark.parametrize("k", [3, 2])
def test_conv_layer_many_tiles(handle, precision, k):
    """More output tiles than compute units: exercises the persistent tile loop
    (and the prefetch of the next tile's input and weights)."""
    rng = np.random.default_rng(15 + k)
    C = K = 100
    H, W = 4 * 70 + k - 1 + 3, 64 * 4 + k - 1 + 5  # partial tiles on both edges
    x = rng.uniform(-1, 1, size=(C, H, W)).astype(np.float32)
    w = glorot(rng, K, C, k, k)
    b = rng.uniform(-0.1, 0.1, size=K).astype(np.float32)
    dt = torch.bfloat16 if precision == "bf16" else torch.float32
    layer = ops.ConvLayer(handle, w, b, precision)
    y = layer(handle, to_hwc(x, capi.feature_stride(C), dt))
    if precision == "bf16":
        x, w = bf16_round(x), bf16_round(w)
    ref = O.conv2d(x, w, b, True)
    np.testing.assert_allclose(from_hwc(y, K), ref, atol=8e-3 if precision == "bf16" else 2e-4, rtol=0)


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
@pytest.mark.parametrize("C", [1, 3])
def test_conv_first_layer_nchw(handle, precision, C):
    rng = np.random.default_rng(20 + C)
    x = rng.normal(size=(C, 11, 45)).astype(np.float32)
    w = glorot(rng, 100, C, 3, 3)
    b = rng.uniform(-0.1, 0.1, size=100).astype(np.float32)
    layer = ops.ConvLayer(handle, w, b, precision)
    y = layer(handle, torch.as_tensor(x, device="cuda"), layout="nchw")
    ref = O.conv2d(x, w, b, True)
    tol = 8e-3 if precision == "bf16" else 2e-5
    np.testing.assert_allclose(from_hwc(y, 100), ref, atol=tol, rtol=0)


@pytest.mark.parametrize("C,K,k,H,W", [(8, 8, 3, 10, 37), (8, 8, 2, 10, 37), (100, 100, 2, 10, 37),
                                       (3, 8, 3, 10, 37), (100, 100, 3, 40, 150), (1, 100, 3, 12, 70),
                                       (64, 64, 3, 9, 66), (20, 40, 3, 7, 20)])
def test_conv_backward(handle, C, K, k, H, W):
    """The adjoint against the oracle's; the weight gradient runs on MFMA
    (several 64-pixel segments per row and per block at 40 x 150)."""
    rng = np.random.default_rng(30 + C + k)
    x = rng.uniform(-1, 1, size=(C, H, W)).astype(np.float32)
    w = glorot(rng, K, C, k, k)
    b = rng.uniform(-0.1, 0.1, size=K).astype(np.float32)
    Ho, Wo = H - k + 1, W - k + 1
    dy = rng.normal(size=(K, Ho, Wo)).astype(np.float32)
    xg = to_hwc(x, capi.feature_stride(C))
    dyg = to_hwc(dy, capi.feature_stride(K))
    dx, dw, db = ops.conv2d_backward(handle, xg, torch.as_tensor(w, device="cuda"),
                                     torch.as_tensor(b, device="cuda"), dyg)
    rdx, rdw, rdb = O.conv2d_bwd(x, w, b, dy, True)
    np.testing.assert_allclose(from_hwc(dx, C), rdx, atol=2e-4 * np.abs(rdx).max(), rtol=0)
    np.testing.assert_allclose(dw.cpu().numpy(), rdw, atol=2e-4 * np.abs(rdw).max(), rtol=0)
    np.testing.assert_allclose(db.cpu().numpy(), rdb, atol=2e-4 * np.abs(rdb).max(), rtol=0)


@pytest.mark.parametrize("C,K,k,x_bf16", [(100, 100, 3, True), (100, 100, 2, True), (1, 100, 3, False)])
def test_conv_backward_stored_bf16(handle, C, K, k, x_bf16):
    """BF16 form of the adjoint from the kept output: against the oracle's
    adjoint of the bf16-rounded operands (x, w, dy), to the bf16 rounding of
    the kept output y, of dz and of dx (1e-2 relative L2)."""
    rng = np.random.default_rng(35 + C + k)
    H, W = 14, 90
    x = rng.uniform(-1, 1, size=(C, H, W)).astype(np.float32)
    w = glorot(rng, K, C, k, k)
    b = rng.uniform(-0.1, 0.1, size=K).astype(np.float32)
    xg = to_hwc(x, capi.feature_stride(C))
    xin = xg.to(torch.bfloat16) if x_bf16 else xg
    lay = ops.ConvLayer(handle, w, b, "bf16")
    y = lay(handle, xin) if x_bf16 else torch.empty(0)
    if not x_bf16:  # a float32 planar image under a bf16 layer, as the towers' first layer reads it
        img = torch.as_tensor(x, device="cuda")
        y = torch.empty((H - k + 1, W - k + 1, capi.feature_stride(K)), dtype=torch.bfloat16, device="cuda")
        call("ccs_conv2d_forward", handle.h, ptr(img), capi.LAYOUT_NCHW, H, W, C, ptr(lay.w), ptr(lay.b), K, k, k,
             1, capi.BF16, ptr(y))
    Ho, Wo = H - k + 1, W - k + 1
    dy = rng.normal(size=(K, Ho, Wo)).astype(np.float32)
    dyg = to_hwc(dy, capi.feature_stride(K)).to(torch.bfloat16)
    dx, dw, db = ops.conv2d_backward_stored(handle, xin, y, lay.w, dyg, True, want_dx=x_bf16)
    xr = x if not x_bf16 else bf16_round(x)
    rdx, rdw, rdb = O.conv2d_bwd(xr, bf16_round(w), b, bf16_round(dy), True)

    def rel(a, r):
        return np.linalg.norm(a - r) / np.linalg.norm(r)

    assert rel(dw.cpu().numpy(), rdw) < 1e-2
    assert rel(db.cpu().numpy(), rdb) < 1e-2
    if x_bf16:
        assert rel(from_hwc(dx, C), rdx) < 1e-2


@pytest.mark.parametrize("k,H,W", [(3, 9, 70), (3, 21, 150), (2, 16, 131), (3, 66, 300), (2, 5, 40)])
def test_dw_bf16_kernel_matches_fp32_mfma(handle, k, H, W):
    """The bf16 weight-gradient kernel (bf16 input and dz, tap rows on
    v_mfma_f32_16x16x32_bf16) against the float32-MFMA kernel on the same values
    (the input given as float32): the products are exact in float32, so dW and
    db agree to the summation order (1e-5 relative, per tap and channel pair
    within 1e-5 of the largest entry).  Widths that end inside a 64-pixel
    segment and heights of one to a few segments per chunk are included."""
    rng = np.random.default_rng(70 + k + H + W)
    C = K = 100
    Ho, Wo = H - k + 1, W - k + 1
    cs = capi.feature_stride(C)
    x = bf16_round(rng.uniform(-1, 1, size=(C, H, W)).astype(np.float32))
    y = bf16_round(rng.uniform(-0.95, 0.95, size=(K, Ho, Wo)).astype(np.float32))
    dy = bf16_round(rng.normal(size=(K, Ho, Wo)).astype(np.float32))
    w = glorot(rng, K, C, k, k)
    wd = torch.as_tensor(w, device="cuda")
    xb = to_hwc(x, cs).to(torch.bfloat16)
    xf = to_hwc(x, cs)
    yb = to_hwc(y, cs).to(torch.bfloat16)
    dyb = to_hwc(dy, cs).to(torch.bfloat16)
    out = []
    for xin, xprec in ((xb, capi.BF16), (xf, capi.FP32)):
        dw = torch.empty((K, C, k, k), device="cuda")
        db = torch.empty(K, device="cuda")
        call("ccs_conv2d_backward_stored", handle.h, ptr(xin), xprec, ptr(yb), capi.BF16, H, W, C, ptr(wd), K, k, k,
             1, ptr(dyb), None, ptr(dw), ptr(db))
        out.append((dw.cpu().numpy().astype(np.float64), db.cpu().numpy().astype(np.float64)))
    (dw_b, db_b), (dw_f, db_f) = out
    scale = np.abs(dw_f).max()
    assert np.linalg.norm(dw_b - dw_f) <= 1e-5 * np.linalg.norm(dw_f)
    assert np.abs(dw_b - dw_f).max() <= 1e-5 * scale
    np.testing.assert_array_equal(db_b, db_f)  # the same bias kernel


# ---------------------------------------------------------------- correlation

def offsets(dmin, dstep, D):
    return -(dmin + np.arange(D) * dstep)


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
@pytest.mark.parametrize("D,dmin,dstep,C", [(32, 0, 1, 100), (128, 0, 1, 100), (16, 3, 2, 8),
                                            (64, -63, 1, 100), (256, 0, 1, 8), (12, 1, 1, 5)])
def test_correlation_volume(handle, precision, D, dmin, dstep, C):
    rng = np.random.default_rng(40 + D + C)
    H, W = 3, 150
    f0 = np.tanh(rng.normal(size=(C, H, W))).astype(np.float32)
    f1 = np.tanh(rng.normal(size=(C, H, W))).astype(np.float32)
    dt = torch.bfloat16 if precision == "bf16" else torch.float32
    cs = capi.feature_stride(C)
    g0, g1 = to_hwc(f0, cs, dt), to_hwc(f1, cs, dt)
    raw = ops.correlation_volume(handle, g0, g1, C, dmin, dstep, D, epilogue="raw")
    cost = ops.correlation_volume(handle, g0, g1, C, dmin, dstep, D, epilogue="neg_softmax")
    if precision == "bf16":
        f0, f1 = bf16_round(f0), bf16_round(f1)
    S = O.correlation(f0, f1, offsets(dmin, dstep, D))
    scale = np.abs(S).max()
    tol = 1e-5 if precision == "fp32" else 1e-5
    np.testing.assert_allclose(raw.cpu().numpy(), S, atol=tol * scale + 1e-5, rtol=0)
    np.testing.assert_allclose(cost.cpu().numpy(), O.neg_softmax(S), atol=2e-6, rtol=1e-4)


@pytest.mark.parametrize("layout", ["nhwc", "nchw"])
def test_correlation_static_compat(handle, layout):
    """StereoCorrelation::forward/backward statics on N x C x H x W float32."""
    rng = np.random.default_rng(50)
    N, C, H, W, D = 1, 7, 5, 40, 9
    f0 = rng.normal(size=(N, C, H, W)).astype(np.float32)
    f1 = rng.normal(size=(N, C, H, W)).astype(np.float32)
    disp = (-np.arange(D)).astype(np.float32)
    t0, t1 = torch.as_tensor(f0, device="cuda"), torch.as_tensor(f1, device="cuda")
    td = torch.as_tensor(disp, device="cuda")
    out = ops.correlation_forward_nchw(handle, t0, t1, td, layout=layout)
    S = O.correlation(f0[0], f1[0], disp.astype(int))
    ref = S[None] if layout == "nhwc" else S.transpose(2, 0, 1)[None]
    np.testing.assert_allclose(out.cpu().numpy(), ref, atol=1e-4)
    G = rng.normal(size=ref.shape).astype(np.float32)
    d0, d1 = ops.correlation_backward_nchw(handle, torch.as_tensor(G, device="cuda"), t0, t1, td, layout=layout)
    Gh = G[0] if layout == "nhwc" else G[0].transpose(1, 2, 0)
    r0, r1 = O.correlation_bwd(f0[0], f1[0], disp.astype(int), Gh)
    np.testing.assert_allclose(d0.cpu().numpy()[0], r0, atol=1e-4)
    np.testing.assert_allclose(d1.cpu().numpy()[0], r1, atol=1e-4)


def test_correlation_static_rect_corr(handle):
    rng = np.random.default_rng(51)
    f0 = rng.normal(size=(1, 4, 6, 30)).astype(np.float32)
    f1 = rng.normal(size=(1, 4, 6, 30)).astype(np.float32)
    disp = (-np.arange(5)).astype(np.float32)
    out = ops.correlation_forward_nchw(handle, torch.as_tensor(f0, device="cuda"), torch.as_tensor(f1, device="cuda"),
                                       torch.as_tensor(disp, device="cuda"), rect_corr=1)
    ref = O.correlation(f0[0], f1[0], disp.astype(int), rect_corr=1)
    np.testing.assert_allclose(out.cpu().numpy()[0], ref, atol=1e-4)


def test_correlation_volume_backward(handle):
    rng = np.random.default_rng(52)
    C, H, W, D = 8, 4, 45, 12
    f0 = rng.normal(size=(C, H, W)).astype(np.float32)
    f1 = rng.normal(size=(C, H, W)).astype(np.float32)
    G = rng.normal(size=(H, W, D)).astype(np.float32)
    d0, d1 = ops.correlation_volume_backward(handle, to_hwc(f0, 8), to_hwc(f1, 8), C, 2, 1, D,
                                             torch.as_tensor(G, device="cuda"))
    r0, r1 = O.correlation_bwd(f0, f1, offsets(2, 1, D), G)
    np.testing.assert_allclose(from_hwc(d0, C), r0, atol=1e-4)
    np.testing.assert_allclose(from_hwc(d1, C), r1, atol=1e-4)


def test_softmax_xent(handle):
    rng = np.random.default_rng(53)
    S = rng.normal(size=(5, 7, 24)).astype(np.float32) * 3
    t = rng.integers(-1, 24, size=(5, 7)).astype(np.int32)
    loss, dS = ops.softmax_xent(handle, torch.as_tensor(S, device="cuda"), torch.as_tensor(t, device="cuda"))
    rl, rdS = O.softmax_xent(S, t)
    assert abs(loss.item() - rl) <= 1e-5 * abs(rl)
    np.testing.assert_allclose(dS.cpu().numpy(), rdS, atol=1e-6)


# ---------------------------------------------------------------- decode

def test_argmin_ties_and_remap(handle):
    rng = np.random.default_rng(60)
    vol = rng.integers(0, 5, size=(9, 13, 37)).astype(np.float32)  # many ties
    out = ops.argmin(handle, torch.as_tensor(vol, device="cuda"), 2.0, 0.5)
    idx = O.argmin(vol)
    np.testing.assert_array_equal(out.cpu().numpy(), 2.0 + 0.5 * idx)
    vol2 = rng.normal(size=(4, 6, 200)).astype(np.float32)
    out2 = ops.argmin(handle, torch.as_tensor(vol2, device="cuda"))
    np.testing.assert_array_equal(out2.cpu().numpy(), O.argmin(vol2).astype(np.float32))
    lab = rng.integers(0, 64, size=(7, 11)).astype(np.int32)
    r = ops.remap(handle, torch.as_tensor(lab, device="cuda"), -63.0, 1.0)
    np.testing.assert_array_equal(r.cpu().numpy(), O.remap(lab, -63.0, 1.0).astype(np.float32))


# ---------------------------------------------------------------- CRF

def crf_case(seed, H, W, D, sharp=5.0):
    rng = np.random.default_rng(seed)
    C = O.neg_softmax(rng.normal(size=(H, W, D)) * sharp).astype(np.float32)
    wx = rng.uniform(0.2, 1.0, size=(H, W)).astype(np.float32)
    wy = rng.uniform(0.2, 1.0, size=(H, W)).astype(np.float32)
    return C, wx, wy


def check_crf(handle, C, wx, wy, L1, L2, delta, iters, gap_eps=1e-3):
    g = ops.slackprop(handle, torch.as_tensor(C, device="cuda"), torch.as_tensor(wx, device="cuda"),
                      torch.as_tensor(wy, device="cuda"), L1, L2, delta, iters)
    r = O.slackprop(C, wx, wy, L1, L2, delta, iters)
    sums = g["sums"].cpu().numpy()
    S_g, K_g = sums[:2 * iters + 1], sums[2 * iters + 1:]
    scale = max(1.0, np.abs(r["sums"]).max())
    np.testing.assert_allclose(S_g, r["sums"], atol=1e-5 * scale, rtol=0)
    if iters:
        np.testing.assert_allclose(K_g, r["kept"], atol=1e-5 * scale, rtol=0)
    assert abs(g["lb"].item() - r["lb"]) <= 1e-5 * max(1.0, abs(r["lb"]))
    lab_g = g["labels"].cpu().numpy()
    mism = lab_g != r["labels"]
    assert np.all(r["gap"][mism] < gap_eps), "label mismatch where the oracle's min-marginal gap is large"
    e_g = O.energy(C, wx, wy, lab_g, L1, L2, delta)
    e_r = O.energy(C, wx, wy, r["labels"], L1, L2, delta)
    assert abs(e_g - e_r) <= 1e-5 * max(1.0, abs(e_r))
    # The C ABI takes L1/L2 as float32: evaluate the oracle energy with the same values.
    e_g32 = O.energy(C, wx, wy, lab_g, float(np.float32(L1)), float(np.float32(L2)), delta)
    eg = ops.crf_energy(handle, torch.as_tensor(C, device="cuda"), torch.as_tensor(wx, device="cuda"),
                        torch.as_tensor(wy, device="cuda"), g["labels"], L1, L2, delta)
    assert abs(eg.item() - e_g32) <= 1e-9 * max(1.0, abs(e_g32))
    return g, r, mism.mean()


@pytest.mark.parametrize("H,W,D", [(6, 40, 16), (5, 33, 128), (4, 21, 200), (1, 50, 32), (30, 1, 32), (7, 9, 64)])
def test_slackprop_vs_oracle(handle, H, W, D):
    C, wx, wy = crf_case(70 + D, H, W, D)
    check_crf(handle, C, wx, wy, 0.5, 2.0, 2, 3)


@pytest.mark.parametrize("L1,L2,delta,iters", [(0.0, 0.0, 2, 2), (0.3, 0.3, 1, 2), (1.0, 12.0, 3, 4),
                                              (0.5, 2.0, 2, 0)])
def test_slackprop_parameters(handle, L1, L2, delta, iters):
    C, wx, wy = crf_case(80, 8, 24, 32)
    check_crf(handle, C, wx, wy, L1, L2, delta, iters)


def test_slackprop_static_and_bad_args(handle):
    C, wx, wy = crf_case(81, 5, 17, 16)
    out, lb = ops.slackprop_forward_static(handle, torch.as_tensor(C, device="cuda"),
                                           torch.as_tensor(wx, device="cuda"), torch.as_tensor(wy, device="cuda"),
                                           0.5, 2.0, 3)
    r = O.slackprop(C, wx, wy, 0.5, 2.0, 2, 3)
    assert abs(lb - r["lb"]) <= 1e-5 * max(1.0, abs(r["lb"]))
    mism = out.cpu().numpy().astype(int) != r["labels"]
    assert np.all(r["gap"][mism] < 1e-3)
    with pytest.raises(capi.CCSError):
        ops.slackprop(handle, torch.as_tensor(C, device="cuda"), torch.as_tensor(wx, device="cuda"),
                      torch.as_tensor(wy, device="cuda"), 2.0, 1.0, 2, 1)  # L2 < L1


def test_ssvm_grad(handle):
    rng = np.random.default_rng(90)
    H, W, D = 6, 11, 9
    xh = rng.integers(0, D, size=(H, W)).astype(np.int32)
    xs = rng.integers(0, D, size=(H, W)).astype(np.int32)
    wx = rng.uniform(0.1, 1, size=(H, W)).astype(np.float32)
    wy = rng.uniform(0.1, 1, size=(H, W)).astype(np.float32)
    dC, dL, dwx, dwy = ops.crf_ssvm_grad(handle, torch.as_tensor(xh, device="cuda"), torch.as_tensor(xs, device="cuda"),
                                         torch.as_tensor(wx, device="cuda"), torch.as_tensor(wy, device="cuda"),
                                         D, 0.5, 2.0, 2)
    rdC, rdL, rdwx, rdwy = O.ssvm_grad(xh, xs, wx, wy, D, 0.5, 2.0, 2)
    np.testing.assert_array_equal(dC.cpu().numpy(), rdC)
    np.testing.assert_allclose(dL.cpu().numpy(), rdL, rtol=1e-12, atol=1e-9)
    np.testing.assert_array_equal(dwx.cpu().numpy(), rdwx)
    np.testing.assert_array_equal(dwy.cpu().numpy(), rdwy)


# ---------------------------------------------------------- refinement, LR check

def test_quad_fit_direct(handle):
    rng = np.random.default_rng(71)
    H, W, D, dmin, step = 9, 23, 32, 1.0, 0.5
    z = np.arange(D)[None, None, :]
    lab = rng.integers(0, D, (H, W))
    lab[0, :3] = [0, D - 1, D - 2]  # both ends and the last refinable label
    # a noisy parabola whose vertex lies within half a label of the pixel's label,
    # so the three values the fit reads are O(1) (well conditioned in float32)
    centre = lab[..., None] + rng.uniform(-0.5, 0.5, (H, W, 1))
    vol = (rng.uniform(0.2, 2.0, (H, W, 1)) * (z - centre) ** 2 + rng.normal(0, 0.05, (H, W, D))).astype(np.float32)
    disp = (dmin + step * lab).astype(np.float32)
    d = torch.as_tensor(disp, device="cuda").clone()
    v = torch.as_tensor(vol, device="cuda")
    capi.call("ccs_quad_fit_direct", handle.h, capi.ptr(d), capi.ptr(v), H, W, D, dmin, step)
    want = O.quad_fit(disp.astype(np.float64), vol.astype(np.float64), dmin, step)
    got = d.cpu().numpy()
    np.testing.assert_allclose(got, want, rtol=0, atol=2e-5)
    assert got[0, 0] == disp[0, 0] and got[0, 1] == disp[0, 1]  # z = 0 and z = D - 1 keep d
    assert got[0, 2] != disp[0, 2]


def test_left_right_check(handle):
    rng = np.random.default_rng(72)
    H, W, th = 11, 57, 1.0
    left = rng.uniform(0, 20, (H, W)).astype(np.float32)
    right = (-rng.uniform(0, 20, (H, W))).astype(np.float32)
    right[:, 10:30] = -left[:, 10:30]  # a consistent band (up to the interpolation)
    out = torch.empty(H, W, device="cuda")
    lt, rt = torch.as_tensor(left, device="cuda"), torch.as_tensor(right, device="cuda")
    capi.call("ccs_left_right_check", handle.h, capi.ptr(lt), capi.ptr(rt), H, W, th, capi.ptr(out))
    got = out.cpu().numpy()
    want = O.lr_check(left.astype(np.float64), right.astype(np.float64), th)
    # the float32 kernel and the float64 oracle may disagree only at |I1 - I2| within 1e-4 of th
    xs = np.arange(W)[None, :] - left.astype(np.float64)
    i0 = np.clip(np.floor(xs), 0, W - 1).astype(int)
    i1 = np.clip(np.floor(xs) + 1, 0, W - 1).astype(int)
    a = xs - np.floor(xs)
    rows = np.arange(H)[:, None]
    i2 = -((1 - a) * right[rows, i0] + a * right[rows, i1])
    clear = np.abs(np.abs(left - i2) - th) > 1e-4
    np.testing.assert_array_equal(got[clear], want[clear])
    assert np.all((got == left) | (got == -1.0))
    assert (got != -1.0).sum() > 0 and (got == -1.0).sum() > 0
