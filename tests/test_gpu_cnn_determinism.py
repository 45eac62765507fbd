"""Run-to-run determinism of the BF16 tower layers and of both correlation
epilogues at the 1242 x 375, D = 128 frame size.

The chain sweeps of the CRF once gave results that changed from call to call
when they stored 128-bit vectors (see tests/test_gpu_crf_determinism.py).  The
conv kernel stores its bf16 outputs as 128-bit vectors and the correlation its
volume as float4, so each of them runs on two fresh handles (the first call of
a handle, which sizes its scratch and packs the weights, included) and again on
the same handle, and every output must be bit-identical.
"""
import numpy as np
import pytest
import torch

from cnn_crf_stereo_amd import capi, ops, synth

pytestmark = pytest.mark.gpu

H, W, D = 375, 1242, 128


def frame_outputs(h, x0, params):
    """Layer 1 (FP32 image -> bf16), a 3x3 100->100 layer, the last layer of a
    7-layer tower, and the raw and -softmax correlation volumes."""
    tower = ops.Tower(h, params, "bf16")
    feats = tower(h, x0)
    lay = tower.layers[1]
    y1 = tower.layers[0](h, x0, layout="nchw")
    y2 = lay(h, y1)
    raw = ops.correlation_volume(h, feats, feats.flip(0).contiguous(), 100, 0, 1, D, epilogue="raw")
    neg = ops.correlation_volume(h, feats, feats.flip(0).contiguous(), 100, 0, 1, D)
    torch.cuda.synchronize()
    return [y1, y2, feats, raw, neg]


@pytest.mark.parametrize("arch", ["all3x3", "reference"])
def test_tower_and_correlation_bit_identical_across_handles_and_calls(arch):
    data = synth.stereo_pair(dict(W=W, H=H, C=1, D=D, planes=12, iters=8), seed=77)
    layers, pad = synth.architecture(7, 1, arch)
    params = synth.glorot_weights(layers, seed=78)
    runs = []
    for _ in range(2):
        h = capi.Handle()
        x0 = ops.pad_normalize(h, torch.as_tensor(data["left"], device="cuda"), pad)
        runs.append(frame_outputs(h, x0, params))  # the handle's first calls
        runs.append(frame_outputs(h, x0, params))
        h.close()
    ref = runs[0]
    assert bool(torch.isfinite(ref[4]).all())
    for other in runs[1:]:
        for a, b in zip(ref, other):
            assert torch.equal(a.view(torch.int16) if a.dtype == torch.bfloat16 else a.view(torch.int32),
                               b.view(torch.int16) if b.dtype == torch.bfloat16 else b.view(torch.int32))
    assert np.isclose(float(ref[4].sum()), -H * W, rtol=1e-4)  # -softmax rows sum to -1
