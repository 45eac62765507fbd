"""Training steps at the sizes of the differentiation workloads.

* 4 x (256 x 512) frames, D = 128, 7 all-3x3 layers, 8 CRF iterations: the
  checkpointed FP32 towers (padded image kept, layer outputs recomputed in the
  backward) give the gradients of the store-everything FP32 layers to 1e-5
  relative (L2 over every parameter, and per tensor); one SGD step lowers the
  loss on the batch.
* 1488 x 992, D = 256, 10 CRF iterations: the checkpointed BF16 step (BF16
  MFMA towers and correlation from FP32 master weights, FP32 softmax, CRF and
  accumulation) against the FP32 store-everything backward at the same
  weights, and its peak device allocation against the store-everything one.

Both go through the torch.autograd.Function harness (cnn-crf-stereo_amd/
autograd.py) and the hand-written adjoints of libccs.so; the adjoints
themselves are pinned against FP64 finite differences and the oracle in
tests/test_gpu_autograd.py.
"""
import numpy as np
import pytest
import torch

from cnn_crf_stereo_amd import autograd as AG
from cnn_crf_stereo_amd import synth

pytestmark = pytest.mark.gpu


def frames(workload, n):
    spec = dict(synth.WORKLOADS[workload])
    out = []
    for i in range(n):
        d = synth.stereo_pair(spec, seed=3000 + i)
        tgt = np.clip(np.rint(d["disp"]), 0, spec["D"] - 1).astype(np.int32)
        tgt[:, :8] = -1  # no ground truth along the left border
        out.append((torch.as_tensor(d["left"], device="cuda"), torch.as_tensor(d["right"], device="cuda"),
                    torch.as_tensor(tgt, device="cuda")))
    return spec, out


def rel_errors(g, ref):
    num = sum(float(((a - b).double() ** 2).sum()) for a, b in zip(g, ref))
    den = sum(float((b.double() ** 2).sum()) for b in ref)
    per = [float((a - b).double().norm() / max(float(b.double().norm()), 1e-30)) for a, b in zip(g, ref)]
    return np.sqrt(num / den), max(per)


def test_batch_256x512_checkpointed_equals_store_everything(handle):
    spec, batch = frames("rds-512x256-d128", 4)
    layers, _ = synth.architecture(7, 1, "all3x3")
    params = synth.glorot_weights(layers, seed=21)
    D = spec["D"]
    plain = AG.StereoModel(params, checkpoint=False)
    loss_p, g_p = AG.gradients(plain, handle, batch, D, iters=8)
    del plain
    ck = AG.StereoModel(params, checkpoint=True)
    loss_c, g_c = AG.gradients(ck, handle, batch, D, iters=8)
    assert np.isfinite(loss_c) and abs(loss_c - loss_p) <= 1e-5 * abs(loss_p)
    tot, worst = rel_errors(g_c, g_p)
    print("checkpointed vs store-everything: rel L2 %.3g, worst tensor %.3g" % (tot, worst))
    assert tot <= 1e-5 and worst <= 1e-5
    loss0 = AG.training_step(ck, handle, batch, D, lr=1e-3, iters=8)
    loss1, _ = AG.gradients(ck, handle, batch, D, iters=8)
    assert loss1 < loss0


def test_1488x992_bf16_step_against_fp32_backward(handle):
    spec, batch = frames("rds-1488x992-d256", 1)
    layers, _ = synth.architecture(7, 1, "all3x3")
    params = synth.glorot_weights(layers, seed=22)
    D = spec["D"]
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    torch.cuda.reset_peak_memory_stats()
    plain = AG.StereoModel(params, checkpoint=False)
    loss_p, g_p = AG.gradients(plain, handle, batch, D, iters=10)
    peak_plain = torch.cuda.max_memory_allocated()
    del plain
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    torch.cuda.reset_peak_memory_stats()
    model = AG.StereoModel(params, precision="bf16", checkpoint=True)
    loss_b, g_b = AG.gradients(model, handle, batch, D, iters=10)
    peak_bf16 = torch.cuda.max_memory_allocated()
    tot, worst = rel_errors(g_b, g_p)
    print("bf16 checkpointed vs fp32 store-everything: loss %.6f / %.6f, grad rel L2 %.3g, worst tensor %.3g, "
          "peak %.2f / %.2f GB" % (loss_b, loss_p, tot, worst, peak_bf16 / 1e9, peak_plain / 1e9))
    assert abs(loss_b - loss_p) <= 1e-2 * abs(loss_p)
    assert tot <= 3e-2 and worst <= 1e-1
    assert peak_bf16 <= 0.6 * peak_plain
