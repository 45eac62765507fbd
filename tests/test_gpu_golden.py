"""GPU path against the committed golden fixtures (tests/golden/*.npz).
Tolerances: float32 features and cost volume absolute 1e-4 (O(1) values),
edge weights relative 1e-5, CRF bound and chain sums relative 1e-5, labels
exact except where the fixture's min-marginal gap is below 1e-3."""
import os

import numpy as np
import pytest
import torch

from cnn_crf_stereo_amd import ops

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def load(name):
    with np.load(os.path.join(HERE, "golden", name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def cuda(a):
    return torch.as_tensor(np.ascontiguousarray(a), device="cuda")


def check_labels(labels, f):
    mism = labels != f["labels"]
    assert np.all(f["gap"][mism] < 1e-3), "labels differ at %d decided pixels" % (f["gap"][mism] >= 1e-3).sum()


def test_frame_fixture_fp32(handle):
    f = load("frame_small")
    D, iters, pad = int(f["D"]), int(f["iters"]), int(f["pad"])
    params = [(f["w%d" % i], f["b%d" % i]) for i in range(3)]
    tower = ops.Tower(handle, params, "fp32")
    feats = [tower(handle, ops.pad_normalize(handle, cuda(img), pad)) for img in (f["left"], f["right"])]
    fl = feats[0][:, :, :8].cpu().numpy().transpose(2, 0, 1)
    np.testing.assert_allclose(fl, f["feat_left"], atol=1e-4, rtol=0)
    cost = ops.correlation_volume(handle, feats[0], feats[1], 8, 0, 1, D)
    np.testing.assert_allclose(cost.cpu().numpy(), f["cost"], atol=1e-4, rtol=0)
    wx, wy = ops.edge_weights(handle, cuda(f["left"][0]), 0.7, 10.0, 1.0, 1.0)
    np.testing.assert_allclose(wx.cpu().numpy(), f["wx"], rtol=1e-5, atol=1e-7)
    np.testing.assert_allclose(wy.cpu().numpy(), f["wy"], rtol=1e-5, atol=1e-7)
    # the CRF on the fixture's own float32 inputs
    r = ops.slackprop(handle, cuda(f["cost"].astype(np.float32)), cuda(f["wx"].astype(np.float32)),
                      cuda(f["wy"].astype(np.float32)), float(f["L1"]), float(f["L2"]), int(f["delta"]), iters)
    lb = r["lb"].item()
    assert abs(lb - f["lb"]) <= 1e-5 * max(1.0, abs(f["lb"]))
    sums = r["sums"].cpu().numpy()
    np.testing.assert_allclose(sums[:2 * iters + 1], f["sums"], rtol=1e-5, atol=1e-5)
    np.testing.assert_allclose(sums[2 * iters + 1:], f["kept"], rtol=1e-5, atol=1e-5)
    check_labels(r["labels"].cpu().numpy(), f)
    disp = ops.remap(handle, r["labels"], 0.0, 1.0).cpu().numpy()
    ok = f["gap"] >= 1e-3
    np.testing.assert_array_equal(disp[ok], f["disparity"][ok])


def test_crf_fixture(handle):
    f = load("crf_small")
    iters = int(f["iters"])
    r = ops.slackprop(handle, cuda(f["cost"]), cuda(f["wx"]), cuda(f["wy"]), float(f["L1"]), float(f["L2"]),
                      int(f["delta"]), iters, want_unaries=True)
    assert abs(r["lb"].item() - f["lb"]) <= 1e-5 * max(1.0, abs(f["lb"]))
    np.testing.assert_allclose(r["unaries"].cpu().numpy(), f["A"], atol=1e-4, rtol=1e-5)
    check_labels(r["labels"].cpu().numpy(), f)
    e = ops.crf_energy(handle, cuda(f["cost"]), cuda(f["wx"]), cuda(f["wy"]), cuda(f["labels"]), float(f["L1"]),
                       float(f["L2"]), int(f["delta"]))
    assert abs(float(e) - f["energy"]) <= 1e-5 * max(1.0, abs(f["energy"]))
