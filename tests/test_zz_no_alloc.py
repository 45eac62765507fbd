"""ccs_conv2d_forward_noalloc against ccs_conv2d_forward, through the C ABI.

The allocation-free entry point packs the weights into a fixed per-GPU array
instead of the handle's scratch and runs the same kernels, so its outputs must
equal the existing entry point's bit for bit.  Each case makes two calls in a
row, the second on a smaller input (and, in the mixed case, with a smaller
weight pack, so the array still holds the first call's larger pack).  On a
fresh handle the call also records into a HIP graph, which the allocating
entry point refuses until its scratch is reserved, and the replays reproduce
the eager output.
"""
import numpy as np
import pytest
import torch

from cnn_crf_stereo_amd import capi
from cnn_crf_stereo_amd.capi import call, ptr

pytestmark = pytest.mark.gpu

PREC = {"fp32": (capi.FP32, torch.float32), "bf16": (capi.BF16, torch.bfloat16)}


def layer(rng, K, C, k):
    a = np.sqrt(6.0 / (C * k * k + K * k * k))
    w = torch.as_tensor(rng.uniform(-a, a, (K, C, k, k)).astype(np.float32), device="cuda")
    b = torch.as_tensor(rng.uniform(-0.1, 0.1, K).astype(np.float32), device="cuda")
    return w, b


def features(rng, H, W, C, dtype):
    cs = capi.feature_stride(C)
    x = torch.zeros((H, W, cs), dtype=torch.float32, device="cuda")
    x[..., :C] = torch.as_tensor(rng.uniform(-1, 1, (H, W, C)).astype(np.float32), device="cuda")
    return x.to(dtype)


def run(name, h, x, w, b, prec):
    H, W, cs = x.shape
    K, C, KH, KW = w.shape
    code, dtype = PREC[prec]
    y = torch.full((H - KH + 1, W - KW + 1, capi.feature_stride(K)), 7.0, dtype=dtype, device="cuda")
    call(name, h.h, ptr(x), capi.LAYOUT_NHWC, H, W, C, ptr(w), ptr(b), K, KH, KW, 1, code, ptr(y))
    return y


# (precision, [(H, W, C, K, k) of the first call, of the second call])
CASES = [
    ("bf16", [(48, 70, 100, 100, 3), (20, 33, 100, 100, 3)]),
    ("fp32", [(40, 52, 100, 100, 3), (17, 21, 100, 100, 2)]),
    ("bf16", [(36, 64, 100, 100, 3), (30, 40, 8, 8, 3)]),   # a smaller pack after a larger one
    ("fp32", [(25, 30, 64, 64, 3), (12, 19, 5, 16, 3)]),    # then a shape without an MFMA kernel
]


@pytest.mark.parametrize("prec,calls", CASES)
def test_noalloc_equals_conv2d_forward(handle, prec, calls):
    rng = np.random.default_rng(hash((prec, calls[0])) % 2**32)
    for (H, W, C, K, k) in calls:
        w, b = layer(rng, K, C, k)
        x = features(rng, H, W, C, PREC[prec][1])
        want = run("ccs_conv2d_forward", handle, x, w, b, prec)
        got = run("ccs_conv2d_forward_noalloc", handle, x, w, b, prec)
        torch.cuda.synchronize()
        assert torch.equal(got, want), (H, W, C, K, k)


# The refused capture of the allocating entry point ends with nothing recorded,
# which torch reports as a warning.
@pytest.mark.filterwarnings("ignore:The CUDA Graph is empty")
@pytest.mark.parametrize("prec", ["bf16", "fp32"])
def test_noalloc_records_into_a_graph_on_a_fresh_handle(prec):
    rng = np.random.default_rng(7)
    w, b = layer(rng, 100, 100, 3)
    x = features(rng, 30, 45, 100, PREC[prec][1])
    stream = torch.cuda.Stream()
    h = capi.Handle(stream)
    code, dtype = PREC[prec]
    y = torch.empty((28, 43, capi.feature_stride(100)), dtype=dtype, device="cuda")
    args = (h.h, ptr(x), capi.LAYOUT_NHWC, 30, 45, 100, ptr(w), ptr(b), 100, 3, 3, 1, code, ptr(y))
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with pytest.raises(capi.CCSError):  # the allocating entry point would have to grow its scratch
        with torch.cuda.graph(g, stream=stream):
            call("ccs_conv2d_forward", *args)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=stream):
        call("ccs_conv2d_forward_noalloc", *args)
    torch.cuda.synchronize()
    h2 = capi.Handle()
    want = run("ccs_conv2d_forward", h2, x, w, b, prec)
    other_w, other_b = layer(rng, 100, 100, 3)
    for _ in range(3):
        y.fill_(3.0)
        run("ccs_conv2d_forward_noalloc", h2, x, other_w, other_b, prec)  # repacks the shared array
        torch.cuda.synchronize()
        with torch.cuda.stream(stream):
            g.replay()  # packs its own weights again, then convolves
        torch.cuda.synchronize()
        assert torch.equal(y, want)
    h.close()
    h2.close()
