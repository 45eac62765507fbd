"""The per-edge ("full") CRF interface, SlackProp::forward(costVolume, pwVolume,
output, max_iter) (stereo/operators/slackprop.h:40, slackprop.cpp:71-108),
against the FP64 oracle's restatement (oracle.slackprop_pe):

* the pairwise volume's re-layout (slackprop.cpp:88-92) equals the oracle's,
  and edges breaking 0 <= P1 <= P2 are flagged;
* ccs_slackprop_pw: chain sums and kept constants within 1e-5 of their scale,
  the bound within 1e-5 relative, labels equal except where the oracle's
  min-marginal gap is below 1e-3, at 1-row, 1-column and 2-D grids, D from 16
  to 256, two penalty levels and the Potts form;
* the reference-order static ccs_slackprop_forward_pw: labels as float and the
  bound, and CCS_ERR_ARG for an invalid volume.
"""
import numpy as np
import pytest
import torch

import oracle as O
from cnn_crf_stereo_amd import capi, ops

pytestmark = pytest.mark.gpu


def dev(a):
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float32), device="cuda")


def case(seed, H, W, D, levels=2, sharp=5.0):
    rng = np.random.default_rng(seed)
    C = O.neg_softmax(rng.normal(size=(H, W, D)) * sharp).astype(np.float32)
    p1 = rng.uniform(0.1, 0.8, size=(2, H, W))  # P1x, P1y
    p2 = p1 + rng.uniform(0.2, 2.5, size=(2, H, W))
    pw = np.concatenate([p1, p2]) if levels == 2 else p1  # channels P1x, P1y, P2x, P2y (or Px, Py)
    return C, pw.astype(np.float32)


def check(handle, C, pw, delta, iters):
    pen, bad = ops.pairwise_penalties(handle, dev(pw))
    assert not bad.any()
    g = ops.slackprop_pw(handle, dev(C), pen, delta, iters)
    p1x, p2x, p1y, p2y = O.pw_penalties(pw)
    r = O.slackprop_pe(C, p1x, p2x, p1y, p2y, delta, iters)
    sums = g["sums"].cpu().numpy()
    scale = max(1.0, np.abs(r["sums"]).max())
    np.testing.assert_allclose(sums[:2 * iters + 1], r["sums"], atol=1e-5 * scale, rtol=0)
    if iters:
        np.testing.assert_allclose(sums[2 * iters + 1:], r["kept"], atol=1e-5 * scale, rtol=0)
    assert abs(g["lb"].item() - r["lb"]) <= 1e-5 * max(1.0, abs(r["lb"]))
    lab = g["labels"].cpu().numpy()
    mism = lab != r["labels"]
    assert np.all(r["gap"][mism] < 1e-3), "label mismatch where the oracle's min-marginal gap is large"
    e_g = O.energy_pe(C, p1x, p2x, p1y, p2y, lab, delta)
    e_r = O.energy_pe(C, p1x, p2x, p1y, p2y, r["labels"], delta)
    assert abs(e_g - e_r) <= 1e-5 * max(1.0, abs(e_r))
    assert r["lb"] <= e_r + 1e-6 * max(1.0, abs(e_r))
    return g, r


def test_relayout_and_validation(handle):
    _, pw = case(400, 6, 9, 8)
    pen, bad = ops.pairwise_penalties(handle, dev(pw))
    np.testing.assert_array_equal(pen.cpu().numpy(), O.pw_relayout(pw.astype(np.float64)).astype(np.float32))
    assert not bad.any()
    for (c, y, x, v) in [(0, 2, 3, -0.1), (3, 1, 4, 0.05)]:  # P1x < 0; P2y < P1y
        wrong = pw.copy()
        wrong[c, y, x] = v
        _, bad = ops.pairwise_penalties(handle, dev(wrong))
        assert bad.cpu().numpy()[(y * 9 + x) // 256] == 1
    wrong = pw.copy()
    wrong[0, 2, 8] = -1.0  # the x penalty of the last column belongs to no edge
    _, bad = ops.pairwise_penalties(handle, dev(wrong))
    assert not bad.any()


@pytest.mark.parametrize("H,W,D", [(6, 40, 16), (5, 33, 128), (4, 21, 200), (1, 50, 32), (30, 1, 32),
                                   (7, 9, 64), (3, 70, 256)])
def test_slackprop_pw_vs_oracle(handle, H, W, D):
    C, pw = case(410 + D, H, W, D)
    check(handle, C, pw, 2, 3)


@pytest.mark.parametrize("levels,delta,iters", [(1, 2, 3), (2, 3, 2), (2, 2, 0)])
def test_slackprop_pw_forms(handle, levels, delta, iters):
    C, pw = case(420 + levels + delta, 8, 24, 32, levels=levels)
    check(handle, C, pw, delta, iters)


def test_slackprop_pw_weighted_equals_slackprop(handle):
    """Per-edge penalties w L1, w L2 give the edge-weight form's results (to the
    rounding of the products)."""
    rng = np.random.default_rng(430)
    H, W, D = 9, 40, 64
    C = O.neg_softmax(rng.normal(size=(H, W, D)) * 5).astype(np.float32)
    wx = rng.uniform(0.2, 1.0, size=(H, W)).astype(np.float32)
    wy = rng.uniform(0.2, 1.0, size=(H, W)).astype(np.float32)
    pw = np.stack([wx * 0.5, wy * 0.5, wx * 2.0, wy * 2.0]).astype(np.float32)
    pen, _ = ops.pairwise_penalties(handle, dev(pw))
    g = ops.slackprop_pw(handle, dev(C), pen, 2, 4)
    s = ops.slackprop(handle, dev(C), dev(wx), dev(wy), 0.5, 2.0, 2, 4)
    assert abs(g["lb"].item() - s["lb"].item()) <= 1e-5 * abs(s["lb"].item())
    r = O.slackprop(C, wx, wy, 0.5, 2.0, 2, 4)
    mism = g["labels"].cpu().numpy() != s["labels"].cpu().numpy()
    assert np.all(r["gap"][mism] < 1e-3)


def test_forward_pw_static(handle):
    C, pw = case(440, 10, 30, 32)
    out, lb = ops.slackprop_forward_pw(handle, dev(C), dev(pw), 3)
    p1x, p2x, p1y, p2y = O.pw_penalties(pw)
    r = O.slackprop_pe(C, p1x, p2x, p1y, p2y, 2, 3)
    assert abs(lb - r["lb"]) <= 1e-5 * max(1.0, abs(r["lb"]))
    mism = out.cpu().numpy().astype(int) != r["labels"]
    assert np.all(r["gap"][mism] < 1e-3)
    _, pwp = case(441, 10, 30, 32, levels=1)  # the Potts form
    out, lb = ops.slackprop_forward_pw(handle, dev(C), dev(pwp), 2)
    q1x, q2x, q1y, q2y = O.pw_penalties(pwp)
    rp = O.slackprop_pe(C, q1x, q2x, q1y, q2y, 2, 2)
    assert abs(lb - rp["lb"]) <= 1e-5 * max(1.0, abs(rp["lb"]))
    wrong = pw.copy()
    wrong[2, 4, 7] = wrong[0, 4, 7] - 0.1  # P2x < P1x
    with pytest.raises(capi.CCSError) as e:
        ops.slackprop_forward_pw(handle, dev(C), dev(wrong), 3)
    assert e.value.status == capi.ERR_ARG
