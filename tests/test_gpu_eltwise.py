"""Stand-alone operator forms and layout moves (bias, tanh, softmax, C4 and
NHWC/NCHW conversions) through the C ABI against torch float32 references of
the same op.  Tolerances: exact for moves and bias adds, 2e-6 relative for
tanh and softmax (float32 exp)."""
import numpy as np
import pytest
import torch

from cnn_crf_stereo_amd.capi import BF16, FP32, LAYOUT_NCHW, LAYOUT_NHWC, call, ptr

pytestmark = pytest.mark.gpu


def test_bias_both_layouts(handle):
    g = torch.Generator(device="cuda").manual_seed(1)
    C, H, W, cs = 5, 7, 9, 8
    b = torch.randn(C, device="cuda", generator=g)
    y = torch.randn(C, H, W, device="cuda", generator=g)
    want = y + b[:, None, None]
    call("ccs_bias_forward", handle.h, ptr(y), LAYOUT_NCHW, C, H, W, 0, ptr(b))
    torch.testing.assert_close(y, want, rtol=0, atol=0)
    z = torch.randn(H, W, cs, device="cuda", generator=g)
    want = z.clone()
    want[:, :, :C] += b
    call("ccs_bias_forward", handle.h, ptr(z), LAYOUT_NHWC, C, H, W, cs, ptr(b))
    torch.testing.assert_close(z, want, rtol=0, atol=0)  # pad channels untouched


def test_tanh(handle):
    x = torch.linspace(-12, 12, 10007, device="cuda")
    want = torch.tanh(x)
    call("ccs_tanh_forward", handle.h, ptr(x), x.numel())
    torch.testing.assert_close(x, want, rtol=2e-6, atol=1e-7)


def test_tanh_ulp_and_special_values(handle):
    """The branch-free tanh: within 4 ulp of the float64 tanh of the float32
    input everywhere (dense around the 0.625 switch point, tiny and large
    arguments), odd, and exact at 0, -0 and +-inf."""
    rng = np.random.default_rng(3)
    xs = np.concatenate([np.linspace(-10, 10, 200001), np.linspace(0.6, 0.65, 20001),
                         np.exp(rng.uniform(np.log(1e-30), np.log(50.0), 100000)),
                         [0.0, 0.625, np.nextafter(np.float32(0.625), np.float32(0)), 1e-38, 88.0, 1e4]])
    xs = np.concatenate([xs, -xs]).astype(np.float32)
    y = torch.as_tensor(xs, device="cuda").clone()
    call("ccs_tanh_forward", handle.h, ptr(y), y.numel())
    got = y.cpu().numpy()
    ref = np.tanh(xs.astype(np.float64))
    ulp = np.spacing(np.abs(ref).astype(np.float32)).astype(np.float64)
    ulp = np.maximum(ulp, np.float64(np.finfo(np.float32).smallest_subnormal))
    assert np.max(np.abs(got - ref) / ulp) <= 4.0
    np.testing.assert_array_equal(got[: len(got) // 2], -got[len(got) // 2:])
    sp = torch.tensor([0.0, -0.0, np.inf, -np.inf], device="cuda")
    call("ccs_tanh_forward", handle.h, ptr(sp), 4)
    out = sp.cpu().numpy()
    np.testing.assert_array_equal(out, [0.0, -0.0, 1.0, -1.0])
    assert np.signbit(out[1]) and not np.signbit(out[0])


@pytest.mark.parametrize("D,negate", [(128, 1), (37, 0), (256, 1), (1, 1)])
def test_softmax_channels(handle, D, negate):
    g = torch.Generator(device="cuda").manual_seed(D)
    x = torch.randn(13, 17, D, device="cuda", generator=g) * 6
    want = torch.softmax(x, -1) * (-1 if negate else 1)
    call("ccs_softmax_forward", handle.h, ptr(x), 13 * 17, D, negate)
    torch.testing.assert_close(x, want, rtol=2e-6, atol=1e-7)


def test_layout_moves(handle):
    g = torch.Generator(device="cuda").manual_seed(3)
    C, cs, H, W = 100, 104, 6, 11
    f = torch.randn(H, W, cs, device="cuda", generator=g)
    out = torch.empty(C, H, W, device="cuda")
    call("ccs_nhwc_to_nchw", handle.h, ptr(f), FP32, C, cs, H, W, ptr(out))
    torch.testing.assert_close(out, f[:, :, :C].permute(2, 0, 1), rtol=0, atol=0)
    fb = f.to(torch.bfloat16)
    call("ccs_nhwc_to_nchw", handle.h, ptr(fb), BF16, C, cs, H, W, ptr(out))
    torch.testing.assert_close(out, fb[:, :, :C].float().permute(2, 0, 1), rtol=0, atol=0)
    back = torch.full((H, W, cs), 7.0, device="cuda")
    call("ccs_nchw_to_nhwc", handle.h, ptr(out), C, cs, H, W, ptr(back))
    torch.testing.assert_close(back[:, :, :C], fb[:, :, :C].float(), rtol=0, atol=0)
    assert torch.all(back[:, :, C:] == 0)


def test_c4_to_planar(handle):
    g = torch.Generator(device="cuda").manual_seed(4)
    H, W, pitch = 5, 7, 4 * 8  # padded rows
    img = torch.randn(H, pitch, device="cuda", generator=g)
    out = torch.empty(3, H, W, device="cuda")
    call("ccs_c4_to_planar", handle.h, ptr(img), H, W, pitch, 3, ptr(out))
    want = img[:, :4 * W].reshape(H, W, 4)[:, :, :3].permute(2, 0, 1)
    torch.testing.assert_close(out, want, rtol=0, atol=0)


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
@pytest.mark.parametrize("C,K", [(64, 2), (5, 3), (17, 1)])
def test_pairwise_head(handle, precision, C, K):
    """|1x1 convolution| into K planes, against numpy float64 on the same
    (bf16-rounded) inputs; 1e-5 relative to the sum of |terms|."""
    rng = np.random.default_rng(C + K)
    H, W = 7, 45
    cs = (C + 7) // 8 * 8
    dt = torch.bfloat16 if precision == "bf16" else torch.float32
    x = torch.zeros(H, W, cs, dtype=dt, device="cuda")
    x[:, :, :C] = torch.as_tensor(rng.uniform(-1, 1, (H, W, C)), dtype=torch.float32, device="cuda").to(dt)
    w = torch.as_tensor(rng.uniform(-0.5, 0.5, (K, C)), dtype=torch.float32, device="cuda")
    out = torch.empty(K, H, W, device="cuda")
    call("ccs_pairwise_head", handle.h, ptr(x), BF16 if precision == "bf16" else FP32, H, W, C, ptr(w), K, ptr(out))
    xf = x[:, :, :C].double().cpu().numpy()
    wf = w.double().cpu().numpy()
    want = np.abs(np.einsum("hwc,kc->khw", xf, wf))
    scale = np.einsum("hwc,kc->khw", np.abs(xf), np.abs(wf))
    assert np.all(np.abs(out.cpu().numpy() - want) <= 1e-5 * scale + 1e-7)


def test_scale_forward(handle):
    y = torch.linspace(-3, 3, 1001, device="cuda")
    want = y * 0.37
    call("ccs_scale_forward", handle.h, ptr(y), y.numel(), 0.37)
    torch.testing.assert_close(y, want, rtol=0, atol=0)
