"""Writes the golden fixtures of this directory from the FP64 oracle.

The reference ships no golden vectors for the hot path (its test inputs are
images and real weights without expected outputs), so the fixtures pin this
build's oracle on small seeded cases; tests/test_golden.py re-derives them
with the oracle on the CPU, and the GPU parity tests compare the HIP path
against them.  Run from the repository root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def synth_pair(seed, C, H, W, dmax):
    """Blurred-noise left image and a right image shifted by a two-plane disparity."""
    rng = np.random.default_rng(seed)
    n = rng.uniform(0, 1, (C, H + 8, W + 8 + dmax))
    k = np.exp(-np.arange(-3, 4) ** 2 / 2.0)
    k /= k.sum()
    for ax in (1, 2):
        n = np.apply_along_axis(lambda v: np.convolve(v, k, mode="same"), ax, n)
    base = n[:, 4:4 + H, 4:]
    d = np.where(np.arange(W)[None, :] < W // 2, dmax // 3, (2 * dmax) // 3) + np.zeros((H, 1), dtype=int)
    left = base[:, :, dmax:dmax + W]
    right = np.empty_like(left)
    for y in range(H):
        for x in range(W):
            right[:, y, x] = base[:, y, dmax + x + d[y, min(W - 1, x)] if x + d[y, x] < W else dmax + x]
    right = right + rng.normal(0, 0.01, right.shape)
    lo, hi = left.min(), left.max()
    return ((left - lo) / (hi - lo)).astype(np.float32), ((right - lo) / (hi - lo)).astype(np.float32)


def frame_small():
    import oracle as O

    C, H, W, D, iters = 1, 24, 48, 16, 3
    left, right = synth_pair(11, C, H, W, D)
    rng = np.random.default_rng(12)
    shapes = [(8, C, 3, 3), (8, 8, 2, 2), (8, 8, 2, 2)]  # the reference's 3x3 + 2x2 layers, 8 channels
    params = []
    for (K, Cin, KH, KW) in shapes:
        a = np.sqrt(6.0 / (Cin * KH * KW + K * KH * KW))
        params.append((rng.uniform(-a, a, (K, Cin, KH, KW)).astype(np.float32),
                       rng.uniform(-0.1, 0.1, K).astype(np.float32)))
    pad = 2
    feats = []
    for img in (left, right):
        x, _, _ = O.normalize(O.pad_symmetric(img.astype(np.float64), pad))
        for w, b in params:
            x = O.conv2d(x, w, b, True)
        feats.append(x)
    S = O.correlation(feats[0], feats[1], -np.arange(D))
    cost = O.neg_softmax(S)
    wx, wy = O.edge_weights(left[0].astype(np.float64), 0.7, 10.0, 1.0, 1.0)
    # the CRF runs on the float32-rounded inputs the GPU sees
    c32, wx32, wy32 = cost.astype(np.float32), wx.astype(np.float32), wy.astype(np.float32)
    r = O.slackprop(c32, wx32, wy32, 0.5, 2.0, 2, iters)
    out = dict(left=left, right=right, D=np.int32(D), iters=np.int32(iters), pad=np.int32(pad),
               L1=np.float64(0.5), L2=np.float64(2.0), delta=np.int32(2),
               feat_left=feats[0], cost=cost, wx=wx, wy=wy, labels=r["labels"], lb=np.float64(r["lb"]),
               sums=r["sums"], kept=r["kept"], gap=r["gap"], disparity=O.remap(r["labels"], 0.0, 1.0))
    for i, (w, b) in enumerate(params):
        out["w%d" % i] = w
        out["b%d" % i] = b
    return out


def crf_small():
    import oracle as O
    import scipy.special

    rng = np.random.default_rng(21)
    H, W, D, iters = 9, 14, 10, 5
    cost = (-scipy.special.softmax(rng.standard_normal((H, W, D)) * 2.5, axis=-1)).astype(np.float32)
    wx = rng.uniform(0.1, 1.0, (H, W)).astype(np.float32)
    wy = rng.uniform(0.1, 1.0, (H, W)).astype(np.float32)
    r = O.slackprop(cost, wx, wy, 0.4, 1.7, 3, iters, want_A=True)
    return dict(cost=cost, wx=wx, wy=wy, L1=np.float64(0.4), L2=np.float64(1.7), delta=np.int32(3),
                iters=np.int32(iters), labels=r["labels"], lb=np.float64(r["lb"]), sums=r["sums"], kept=r["kept"],
                gap=r["gap"], A=r["A"], energy=np.float64(O.energy(cost, wx, wy, r["labels"], 0.4, 1.7, 3)))


FIXTURES = {"frame_small": frame_small, "crf_small": crf_small}

if __name__ == "__main__":
    for name, fn in FIXTURES.items():
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, **fn())
        print("wrote", path, os.path.getsize(path), "bytes")
