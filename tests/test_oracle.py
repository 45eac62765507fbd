"""Pins the FP64 oracle (oracle/oracle.cpp) on the CPU.

Each check is independent of the oracle's own code: a library routine that
computes the same quantity (numpy, scipy, torch-CPU), a closed form, an
invariant, or brute force over every labeling of a tiny problem.  A dropped
term, a flipped sign or index, or a transposed operand in the oracle fails at
least one test here.
"""
import itertools

import numpy as np
import pytest
import scipy.ndimage
import scipy.special
import torch
import torch.nn.functional as F

import oracle as O

RNG = np.random.default_rng(12345)


# --- padding and normalisation (stereo/stereo.cu:865-889, stereo/utils.cu:169-287)

@pytest.mark.parametrize("C,H,W,p", [(1, 7, 9, 4), (3, 5, 6, 2), (2, 8, 3, 3), (1, 1, 5, 2)])
def test_pad_symmetric_matches_numpy(C, H, W, p):
    img = RNG.standard_normal((C, H, W))
    got = O.pad_symmetric(img, p)
    want = np.pad(img, ((0, 0), (p, p), (p, p)), mode="symmetric")
    np.testing.assert_array_equal(got, want)


def test_reflect_period_and_edge_repeat():
    # symmetric reflection repeats the edge pixel: -1 -> 0, n -> n-1
    assert O.reflect(-1, 5) == 0 and O.reflect(5, 5) == 4 and O.reflect(-6, 5) == 4
    assert all(O.reflect(i, 7) == O.reflect(i + 14, 7) for i in range(-20, 20))


def test_normalize_population_std():
    img = RNG.uniform(0, 1, (3, 11, 13)) * np.array([1.0, 5.0, 0.1])[:, None, None]
    x, mean, std = O.normalize(img)
    np.testing.assert_allclose(mean, img.reshape(3, -1).mean(1), rtol=1e-13)
    np.testing.assert_allclose(std, img.reshape(3, -1).std(1, ddof=0), rtol=1e-12)
    np.testing.assert_allclose(x.reshape(3, -1).mean(1), 0.0, atol=1e-13)
    np.testing.assert_allclose(x.reshape(3, -1).std(1, ddof=0), 1.0, rtol=1e-12)


def test_normalize_constant_channel_is_not_scaled():
    img = np.full((1, 4, 4), 3.0)
    x, mean, std = O.normalize(img)
    assert mean[0] == 3.0 and std[0] == 0.0
    np.testing.assert_array_equal(x, 0.0)


def test_rgb2gray_weights():
    rgb = RNG.uniform(0, 1, (3, 4, 5))
    g = O.rgb2gray(rgb)
    # known answers: the literal weights of stereo/stereo.cu:861 (they sum to 0.9999)
    e = np.eye(3)[:, :, None, None] * np.ones((1, 1, 1, 1))
    w = [O.rgb2gray(e[i])[0, 0] for i in range(3)]
    assert w == [0.2989, 0.5870, 0.1140]
    # and the map is linear in the three channels
    np.testing.assert_allclose(g, w[0] * rgb[0] + w[1] * rgb[1] + w[2] * rgb[2], rtol=1e-12)


# --- convolution + bias + tanh (stereo/operators/convolution.cpp:104-152)

@pytest.mark.parametrize("C,K,KH,KW,H,W,act", [(1, 4, 3, 3, 9, 11, True), (5, 3, 2, 2, 6, 7, True),
                                               (3, 2, 3, 2, 5, 8, False), (2, 6, 1, 1, 4, 4, True)])
def test_conv_is_true_convolution(C, K, KH, KW, H, W, act):
    x = RNG.standard_normal((C, H, W))
    w = RNG.standard_normal((K, C, KH, KW))
    b = RNG.standard_normal(K)
    got = O.conv2d(x, w, b, act)
    # cross-correlation (torch) with the flipped kernel == convolution
    y = F.conv2d(torch.from_numpy(x)[None], torch.from_numpy(w).flip(2, 3), torch.from_numpy(b))[0]
    if act:
        y = torch.tanh(y)
    np.testing.assert_allclose(got, y.numpy(), rtol=1e-12, atol=1e-12)


def test_conv_one_hot_kernel_shifts():
    # a kernel with a single 1 at (i, j) moves the image by (KH-1-i, KW-1-j)
    x = RNG.standard_normal((1, 6, 7))
    w = np.zeros((1, 1, 3, 3))
    w[0, 0, 0, 2] = 1.0
    got = O.conv2d(x, w, None, False)
    np.testing.assert_allclose(got[0], x[0, 2:6, 0:5], rtol=0, atol=0)


def test_conv_backward_matches_autograd():
    C, K, KH, KW, H, W = 3, 4, 3, 2, 7, 6
    x = RNG.standard_normal((C, H, W))
    w = RNG.standard_normal((K, C, KH, KW)) * 0.3
    b = RNG.standard_normal(K) * 0.1
    dy = RNG.standard_normal((K, H - KH + 1, W - KW + 1))
    dx, dw, db = O.conv2d_bwd(x, w, b, dy, True)
    xt = torch.from_numpy(x).requires_grad_()
    wt = torch.from_numpy(w).requires_grad_()
    bt = torch.from_numpy(b).requires_grad_()
    y = torch.tanh(F.conv2d(xt[None], wt.flip(2, 3), bt))[0]
    y.backward(torch.from_numpy(dy))
    np.testing.assert_allclose(dx, xt.grad.numpy(), rtol=1e-11, atol=1e-12)
    np.testing.assert_allclose(dw, wt.grad.numpy(), rtol=1e-11, atol=1e-12)
    np.testing.assert_allclose(db, bt.grad.numpy(), rtol=1e-11, atol=1e-12)


# --- correlation (stereo/operators/stereocorrelation.cu:111-142, 243-350)

def _corr_brute(f0, f1, disp, rc=0):
    C, H, W = f0.shape
    out = np.zeros((H, W, len(disp)))
    for d, dv in enumerate(disp):
        for y in range(H):
            for x in range(W):
                xs = x + dv
                if not 0 <= xs < W:
                    continue
                best = -np.inf
                for yy in range(max(0, y - rc), min(H, y + rc + 1)):
                    best = max(best, float(f0[:, y, x] @ f1[:, yy, xs]))
                out[y, x, d] = best
    return out


@pytest.mark.parametrize("rc", [0, 1])
def test_correlation_brute_force(rc):
    f0 = RNG.standard_normal((4, 5, 9))
    f1 = RNG.standard_normal((4, 5, 9))
    disp = np.array([0, -1, -3, -8, 2])
    np.testing.assert_allclose(O.correlation(f0, f1, disp, rc), _corr_brute(f0, f1, disp, rc), rtol=1e-12)


def test_correlation_closed_forms():
    f = RNG.standard_normal((3, 4, 6))
    out = O.correlation(f, f, np.array([0, -6, -7]))
    np.testing.assert_allclose(out[..., 0], (f * f).sum(0), rtol=1e-12)  # f0 = f1 at d = 0
    np.testing.assert_array_equal(out[..., 1:], 0.0)  # every x + disp outside the image
    # one-hot features: out = [c0 == c1] exactly
    C, H, W = 4, 2, 8
    c0 = RNG.integers(0, C, (H, W))
    c1 = RNG.integers(0, C, (H, W))
    h0 = np.moveaxis(np.eye(C)[c0], -1, 0)
    h1 = np.moveaxis(np.eye(C)[c1], -1, 0)
    o = O.correlation(h0, h1, np.array([-2]))
    want = np.zeros((H, W))
    want[:, 2:] = (c0[:, 2:] == c1[:, :-2])
    np.testing.assert_array_equal(o[..., 0], want)


def test_correlation_backward_is_adjoint():
    # corr is bilinear: <df0, h> = <G, corr(h, f1)> and <df1, h> = <G, corr(f0, h)>
    f0 = RNG.standard_normal((3, 4, 7))
    f1 = RNG.standard_normal((3, 4, 7))
    disp = np.array([0, -1, -2, -5])
    G = RNG.standard_normal((4, 7, 4))
    df0, df1 = O.correlation_bwd(f0, f1, disp, G)
    for _ in range(3):
        h = RNG.standard_normal(f0.shape)
        np.testing.assert_allclose((df0 * h).sum(), (G * O.correlation(h, f1, disp)).sum(), rtol=1e-11)
        np.testing.assert_allclose((df1 * h).sum(), (G * O.correlation(f0, h, disp)).sum(), rtol=1e-11)


# --- softmax (stereo/operators/softmax.cpp:30-38, stereo/utils.cu:289-292)

def test_neg_softmax_matches_scipy_and_invariants():
    S = RNG.standard_normal((5, 6, 17)) * 10
    c = O.neg_softmax(S)
    np.testing.assert_allclose(c, -scipy.special.softmax(S, axis=-1), rtol=1e-12, atol=1e-300)
    np.testing.assert_allclose(c.sum(-1), -1.0, rtol=1e-13)
    assert c.max() <= 0.0 and c.min() >= -1.0


def test_softmax_xent_matches_torch():
    S = RNG.standard_normal((30, 9)) * 3
    t = RNG.integers(0, 9, 30).astype(np.int32)
    t[[3, 7]] = 9  # out-of-range target: pixel ignored
    loss, dS = O.softmax_xent(S, t)
    St = torch.from_numpy(S).requires_grad_()
    valid = torch.from_numpy(t < 9)
    lt = F.cross_entropy(St[valid], torch.from_numpy(t.astype(np.int64))[valid], reduction="sum")
    lt.backward()
    np.testing.assert_allclose(loss, lt.item(), rtol=1e-12)
    np.testing.assert_allclose(dS, St.grad.numpy(), rtol=1e-11, atol=1e-14)


# --- edge weights (stereo/stereo.cu:1090-1095, 471-550)

def test_gauss_matches_scipy_reflect():
    img = RNG.uniform(0, 1, (13, 17))
    got = O.gauss(img, 0.7)
    # radius ceil(3 sigma) = 3, symmetric border (scipy's "reflect" repeats the edge)
    want = scipy.ndimage.gaussian_filter(img, 0.7, mode="reflect", truncate=3.0 / 0.7)
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-14)


def test_edge_weights_cases():
    img = RNG.uniform(0, 1, (9, 12))
    wx, wy = O.edge_weights(img, 0.7, 10.0, 1.0, 1.0)
    g = scipy.ndimage.gaussian_filter(img, 0.7, mode="reflect", truncate=3.0 / 0.7)
    gx = np.abs(np.diff(g, axis=1, append=g[:, -1:]))
    gy = np.abs(np.diff(g, axis=0, append=g[-1:, :]))
    np.testing.assert_allclose(wx, np.maximum(np.exp(-10 * gx), 1e-6), rtol=1e-12)
    np.testing.assert_allclose(wy, np.maximum(np.exp(-10 * gy), 1e-6), rtol=1e-12)
    # the last column / row couples a pixel with its mirror: weight lambda
    wx, wy = O.edge_weights(img, 0.7, 20.0, 0.8, 2.5)
    np.testing.assert_allclose(wx[:, -1], 2.5, rtol=1e-14)
    np.testing.assert_allclose(wy[-1, :], 2.5, rtol=1e-14)
    # a constant image gives lambda everywhere; a huge edge is clamped at 1e-6
    c = np.full((6, 6), 0.3)
    assert np.all(O.edge_weights(c, 0.7, 20.0, 0.8, 2.5)[0] == 2.5)
    step = np.zeros((6, 8))
    step[:, 4:] = 1e4
    assert O.edge_weights(step, 0.7, 20.0, 0.8, 1.0)[0].min() == 1e-6


# --- argmin / remap (stereo/stereo.cu:348-370, 841-850)

def test_argmin_first_minimum():
    v = RNG.integers(0, 4, (50, 7)).astype(np.float64)  # many ties
    np.testing.assert_array_equal(O.argmin(v), np.argmin(v, axis=-1))


def test_remap_closed_form():
    lab = RNG.integers(0, 64, (5, 6))
    np.testing.assert_allclose(O.remap(lab, 3.0, 0.5), lab * 0.5 + 3.0, rtol=0)


def test_quad_fit_recovers_parabola_vertex():
    """A volume that is an exact parabola around a sub-label disparity d* is
    refined to d* (the 3-point fit of a quadratic is the quadratic); labels at
    the ends and flat volumes stay put."""
    H, W, D, dmin, step = 4, 5, 16, 2.0, 0.5
    z0 = RNG.integers(0, D, (H, W))
    f = RNG.uniform(-0.5, 0.5, (H, W))
    a = RNG.uniform(0.5, 3.0, (H, W))
    z = np.arange(D)[None, None, :]
    vol = a[..., None] * (z - z0[..., None] - f[..., None]) ** 2 + 7.0
    disp = dmin + step * z0
    got = O.quad_fit(disp, vol, dmin, step)
    inner = (z0 > 0) & (z0 < D - 1)
    np.testing.assert_allclose(got[inner], (dmin + step * (z0 + f))[inner], rtol=0, atol=1e-12)
    np.testing.assert_array_equal(got[~inner], disp[~inner])
    np.testing.assert_array_equal(O.quad_fit(disp, np.ones((H, W, D)), dmin, step), disp)


def test_lr_check_closed_forms():
    """Consistent maps pass, maps off by more than th are rejected, a linear
    right map is interpolated exactly, and samples left of the image clamp to
    column 0."""
    H, W = 3, 12
    left = np.full((H, W), 3.0)
    np.testing.assert_array_equal(O.lr_check(left, -left, 0.5), left)
    np.testing.assert_array_equal(O.lr_check(left, -(left + 2.0), 1.0), -np.ones((H, W)))
    xs = np.arange(W, dtype=np.float64)
    right = -np.broadcast_to(0.1 * xs + 1.0, (H, W))          # R(x) = -(0.1 x + 1)
    left = np.broadcast_to(0.25 + 0.3 * xs, (H, W)).copy()    # sample points x - I1 inside the row for x >= 1
    out = O.lr_check(left, right, 10.0)
    u = xs - left[0]
    want_i2 = 0.1 * np.clip(u, 0, W - 1) + 1.0                # clamped linear interpolation of a linear map
    th = np.abs(left[0] - want_i2)
    for t in (0.2, 0.6, 1.5):
        got = O.lr_check(left, right, t)[0]
        np.testing.assert_array_equal(got, np.where(th <= t, left[0], -1.0))
    assert np.all(out == left)


# --- CRF: energy, chains, SlackProp (restated; stereo/stereo.cpp:129-162)

def _rho_table(D, L1, L2, delta):
    d = np.abs(np.arange(D)[:, None] - np.arange(D)[None, :])
    return np.where(d == 0, 0.0, np.where(d < delta, L1, L2))


def _energy_brute(C, wx, wy, lab, L1, L2, delta):
    H, W, D = C.shape
    R = _rho_table(D, L1, L2, delta)
    e = sum(C[y, x, lab[y, x]] for y in range(H) for x in range(W))
    e += sum(wx[y, x] * R[lab[y, x], lab[y, x + 1]] for y in range(H) for x in range(W - 1))
    e += sum(wy[y, x] * R[lab[y, x], lab[y + 1, x]] for y in range(H - 1) for x in range(W))
    return e


def _all_energies(C, wx, wy, L1, L2, delta):
    """Energy of every labeling of a tiny grid (vectorised over labelings)."""
    H, W, D = C.shape
    n = H * W
    labs = np.array(list(itertools.product(range(D), repeat=n))).reshape(-1, H, W)
    R = _rho_table(D, L1, L2, delta)
    yy, xx = np.meshgrid(np.arange(H), np.arange(W), indexing="ij")
    e = C[yy, xx, labs].sum((1, 2))
    e += (wx[:, :-1] * R[labs[:, :, :-1], labs[:, :, 1:]]).sum((1, 2))
    e += (wy[:-1, :] * R[labs[:, :-1, :], labs[:, 1:, :]]).sum((1, 2))
    return labs, e


def test_rho_levels():
    assert [O.rho(d, 0.5, 2.0, 2) for d in (0, 1, -1, 2, -5)] == [0.0, 0.5, 0.5, 2.0, 2.0]
    assert [O.rho(d, 0.5, 2.0, 3) for d in (0, 1, 2, 3)] == [0.0, 0.5, 0.5, 2.0]


def test_energy_matches_brute_force():
    C = RNG.standard_normal((4, 5, 6))
    wx, wy = RNG.uniform(0, 1, (4, 5)), RNG.uniform(0, 1, (4, 5))
    lab = RNG.integers(0, 6, (4, 5))
    np.testing.assert_allclose(O.energy(C, wx, wy, lab, 0.5, 2.0, 2), _energy_brute(C, wx, wy, lab, 0.5, 2.0, 2),
                               rtol=1e-12)


@pytest.mark.parametrize("delta", [2, 3])
def test_chain_viterbi_is_exact(delta):
    n, D = 5, 4
    th = RNG.standard_normal((n, D))
    w = RNG.uniform(0, 1.5, n)
    C = th[None]
    labs, e = _all_energies(C, w[None], np.zeros((1, n)), 0.4, 1.3, delta)
    E, lab = O.chain_viterbi(th, w, 0.4, 1.3, delta)
    np.testing.assert_allclose(E, e.min(), rtol=1e-12)
    np.testing.assert_allclose(_energy_brute(C, w[None], np.zeros((1, n)), lab[None], 0.4, 1.3, delta), E,
                               rtol=1e-12)


@pytest.mark.parametrize("seed", range(4))
def test_chain_minorant_is_tight_minorant(seed):
    rng = np.random.default_rng(seed)
    n, D = 5, 4
    th = rng.standard_normal((n, D)) * 2
    w = rng.uniform(0, 2, n)
    L1, L2, delta = 0.3 + seed * 0.2, 1.5, 2 + seed % 2
    lam, g, kept = O.chain_minorant(th, w, L1, L2, delta)
    labs, e = _all_energies(th[None], w[None], np.zeros((1, n)), L1, L2, delta)
    # Lambda is the chain minimum
    np.testing.assert_allclose(lam, e.min(), rtol=1e-12, atol=1e-12)
    # sum_i g_i(y_i) + kept <= E(y) for every labeling y ...
    gy = g[np.arange(n)[None, :], labs[:, 0, :]].sum(1) + kept
    assert np.all(gy <= e + 1e-10)
    # ... with equality of the minima (tight)
    np.testing.assert_allclose(g.min(1).sum() + kept, lam, rtol=1e-12, atol=1e-12)


def test_slackprop_single_row_is_exact():
    rng = np.random.default_rng(7)
    W, D = 9, 5
    C = rng.standard_normal((1, W, D))
    wx, wy = rng.uniform(0.2, 1, (1, W)), rng.uniform(0.2, 1, (1, W))
    r = O.slackprop(C, wx, wy, 0.5, 2.0, 2, 3)
    E, lab = O.chain_viterbi(C[0], wx[0], 0.5, 2.0, 2)
    np.testing.assert_allclose(r["lb"], E, rtol=1e-12)
    np.testing.assert_allclose(O.energy(C, wx, wy, r["labels"], 0.5, 2.0, 2), E, rtol=1e-12)


def test_slackprop_single_column_is_exact():
    rng = np.random.default_rng(8)
    H, D = 8, 4
    C = rng.standard_normal((H, 1, D))
    wx, wy = rng.uniform(0.2, 1, (H, 1)), rng.uniform(0.2, 1, (H, 1))
    r = O.slackprop(C, wx, wy, 0.7, 1.5, 2, 4)
    E, _ = O.chain_viterbi(C[:, 0], wy[:, 0], 0.7, 1.5, 2)
    np.testing.assert_allclose(r["lb"], E, rtol=1e-10)
    np.testing.assert_allclose(O.energy(C, wx, wy, r["labels"], 0.7, 1.5, 2), E, rtol=1e-10)


def test_slackprop_without_pairwise_is_argmin():
    C = RNG.standard_normal((6, 7, 5))
    wx, wy = RNG.uniform(0, 1, (6, 7)), RNG.uniform(0, 1, (6, 7))
    r = O.slackprop(C, wx, wy, 0.0, 0.0, 2, 2)
    np.testing.assert_array_equal(r["labels"], np.argmin(C, -1))
    np.testing.assert_allclose(r["lb"], C.min(-1).sum(), rtol=1e-12)
    r = O.slackprop(C, np.zeros((6, 7)), np.zeros((6, 7)), 0.5, 2.0, 2, 2)
    np.testing.assert_array_equal(r["labels"], np.argmin(C, -1))


@pytest.mark.parametrize("seed", range(3))
def test_slackprop_bound_brute_force(seed):
    rng = np.random.default_rng(100 + seed)
    H, W, D = 3, 3, 3
    C = rng.standard_normal((H, W, D))
    wx, wy = rng.uniform(0.5, 1.5, (H, W)), rng.uniform(0.5, 1.5, (H, W))
    _, e = _all_energies(C, wx, wy, 0.6, 1.4, 2)
    r = O.slackprop(C, wx, wy, 0.6, 1.4, 2, 6)
    e_lab = O.energy(C, wx, wy, r["labels"], 0.6, 1.4, 2)
    assert r["lb"] <= e.min() + 1e-10 <= e_lab + 2e-10
    seq = O.dual_sequence(r["sums"], r["kept"])
    assert all(b >= a - 1e-10 for a, b in zip(seq, seq[1:]))  # the dual never decreases
    np.testing.assert_allclose(seq[-1], r["lb"], rtol=1e-12)


def test_slackprop_dual_monotone_larger():
    rng = np.random.default_rng(5)
    C = -scipy.special.softmax(rng.standard_normal((12, 15, 8)) * 3, axis=-1)
    wx, wy = rng.uniform(0.1, 1, (12, 15)), rng.uniform(0.1, 1, (12, 15))
    r = O.slackprop(C, wx, wy, 0.5, 2.0, 2, 8)
    seq = O.dual_sequence(r["sums"], r["kept"])
    assert all(b >= a - 1e-9 for a, b in zip(seq, seq[1:]))
    assert r["lb"] <= O.energy(C, wx, wy, r["labels"], 0.5, 2.0, 2) + 1e-9
    # the decode gap is non-negative and the labels minimise the row min-marginals
    assert r["gap"].min() >= 0


def test_ssvm_grad_closed_form():
    rng = np.random.default_rng(9)
    H, W, D = 4, 5, 6
    xh = rng.integers(0, D, (H, W))
    xs = rng.integers(0, D, (H, W))
    wx, wy = rng.uniform(0, 1, (H, W)), rng.uniform(0, 1, (H, W))
    L1, L2 = 0.5, 2.0
    dC, dL, dwx, dwy = O.ssvm_grad(xh, xs, wx, wy, D, L1, L2, 2)
    # subgradient of the margin E(x*) - E(xhat)
    want = np.eye(D)[xs] - np.eye(D)[xh]
    np.testing.assert_array_equal(dC, want)
    # E(x*) - E(xhat) is linear in L1, L2, wx, wy: exact finite differences
    zero = np.zeros((H, W, D))

    def diff(L1_, L2_, wx_, wy_):
        return (O.energy(zero, wx_, wy_, xs, L1_, L2_, 2) - O.energy(zero, wx_, wy_, xh, L1_, L2_, 2))

    base = diff(L1, L2, wx, wy)
    np.testing.assert_allclose(dL[0], diff(L1 + 1, L2, wx, wy) - base, atol=1e-12)
    np.testing.assert_allclose(dL[1], diff(L1, L2 + 1, wx, wy) - base, atol=1e-12)
    e = np.zeros((H, W))
    e[1, 2] = 1.0
    np.testing.assert_allclose(dwx[1, 2], diff(L1, L2, wx + e, wy) - base, atol=1e-12)
    np.testing.assert_allclose(dwy[1, 2], diff(L1, L2, wx, wy + e) - base, atol=1e-12)


def test_masked_energy_equals_energy_with_dropped_terms():
    """or_energy_masked over the labelled pixels equals or_energy of the same
    labelling once the unlabelled pixels' unaries and every edge weight that
    touches them are set to zero; with every pixel labelled it is or_energy."""
    rng = np.random.default_rng(12)
    H, W, D = 5, 6, 7
    C = rng.normal(size=(H, W, D))
    wx, wy = rng.uniform(0.1, 1, (H, W)), rng.uniform(0.1, 1, (H, W))
    lab = rng.integers(0, D, (H, W))
    mask = rng.integers(0, D, (H, W))
    np.testing.assert_allclose(O.energy_masked(C, wx, wy, lab, mask, 0.5, 2.0, 2), O.energy(C, wx, wy, lab, 0.5, 2.0, 2),
                               rtol=1e-14)
    mask[rng.uniform(size=(H, W)) < 0.3] = -1
    mask[0, 0], mask[2, 3] = -1, D  # both sides of the range
    out = mask < 0
    out |= mask >= D
    C0, wx0, wy0 = C.copy(), wx.copy(), wy.copy()
    C0[out] = 0.0
    wx0[out] = 0.0
    wx0[:, :-1][out[:, 1:]] = 0.0  # edge (x, y)-(x+1, y) ends at an unlabelled pixel
    wy0[out] = 0.0
    wy0[:-1, :][out[1:, :]] = 0.0
    np.testing.assert_allclose(O.energy_masked(C, wx, wy, lab, mask, 0.5, 2.0, 2),
                               O.energy(C0, wx0, wy0, lab, 0.5, 2.0, 2), rtol=1e-12)


def test_ssvm_grad_with_unlabelled_pixels():
    """With ground truth missing at some pixels (x* outside [0, D)) the
    subgradient is that of the masked margin E_m(x*) - E_m(xhat): 0 on the
    unlabelled pixels' unaries, and exact finite differences of the masked
    energies in L1, L2 and the edge weights."""
    rng = np.random.default_rng(13)
    H, W, D = 5, 6, 6
    xh = rng.integers(0, D, (H, W))
    xs = rng.integers(0, D, (H, W))
    xs[rng.uniform(size=(H, W)) < 0.3] = -1
    xs[1, 2] = -1
    wx, wy = rng.uniform(0, 1, (H, W)), rng.uniform(0, 1, (H, W))
    L1, L2 = 0.5, 2.0
    dC, dL, dwx, dwy = O.ssvm_grad(xh, xs, wx, wy, D, L1, L2, 2)
    lab = xs >= 0
    want = np.where(lab[..., None], np.eye(D)[np.clip(xs, 0, D - 1)] - np.eye(D)[xh], 0.0)
    np.testing.assert_array_equal(dC, want)
    zero = np.zeros((H, W, D))
    xsc = np.clip(xs, 0, D - 1)

    def diff(L1_, L2_, wx_, wy_):
        return (O.energy_masked(zero, wx_, wy_, xsc, xs, L1_, L2_, 2)
                - O.energy_masked(zero, wx_, wy_, xh, xs, L1_, L2_, 2))

    base = diff(L1, L2, wx, wy)
    np.testing.assert_allclose(dL[0], diff(L1 + 1, L2, wx, wy) - base, atol=1e-12)
    np.testing.assert_allclose(dL[1], diff(L1, L2 + 1, wx, wy) - base, atol=1e-12)
    for (y, x) in [(1, 1), (1, 2), (3, 4)]:
        e = np.zeros((H, W))
        e[y, x] = 1.0
        np.testing.assert_allclose(dwx[y, x], diff(L1, L2, wx + e, wy) - base, atol=1e-12)
        np.testing.assert_allclose(dwy[y, x], diff(L1, L2, wx, wy + e) - base, atol=1e-12)


# ---------------------------------------------------------------- per-edge CRF

def test_pw_relayout_matches_reshape_flip_permute():
    """The pairwise volume's re-layout as the reference writes it with
    ndarray (stereo/operators/slackprop.cpp:88-92): channels split into
    (2, levels) with x/y fastest, levels flipped, x/y moved last."""
    rng = np.random.default_rng(20)
    for levels in (1, 2, 3):
        pw = rng.normal(size=(2 * levels, 4, 5))
        want = np.transpose(pw.reshape(levels, 2, 4, 5)[::-1], (0, 2, 3, 1))
        np.testing.assert_array_equal(O.pw_relayout(pw), want)
    # the reference's channel order P1x, P1y, P2x, P2y: P1 = pen[levels - 1], P2 = pen[0]
    pw = np.stack([np.full((2, 3), v) for v in (1.0, 2.0, 3.0, 4.0)])
    p1x, p2x, p1y, p2y = O.pw_penalties(pw)
    assert (p1x == 1).all() and (p1y == 2).all() and (p2x == 3).all() and (p2y == 4).all()


def test_slackprop_pe_with_weighted_penalties_is_slackprop():
    """Per-edge penalties P1 = w L1, P2 = w L2 reproduce or_slackprop: labels and
    unaries bit for bit, the float64 sums to their summation order (OpenMP
    reductions add the chains in a run-dependent order)."""
    rng = np.random.default_rng(21)
    H, W, D = 5, 7, 6
    C = rng.normal(size=(H, W, D))
    wx, wy = rng.uniform(0.2, 1, (H, W)), rng.uniform(0.2, 1, (H, W))
    r = O.slackprop(C, wx, wy, 0.5, 2.0, 2, 3, want_A=True)
    q = O.slackprop_pe(C, wx * 0.5, wx * 2.0, wy * 0.5, wy * 2.0, 2, 3, want_A=True)
    assert abs(r["lb"] - q["lb"]) <= 1e-12 * abs(r["lb"])
    np.testing.assert_array_equal(r["labels"], q["labels"])
    np.testing.assert_array_equal(r["A"], q["A"])
    np.testing.assert_allclose(r["sums"], q["sums"], rtol=1e-12)


def _brute_min(C, p1x, p2x, p1y, p2y, delta):
    H, W, D = C.shape
    best = np.inf
    for lab in itertools.product(range(D), repeat=H * W):
        e = O.energy_pe(C, p1x, p2x, p1y, p2y, np.array(lab, dtype=np.int32).reshape(H, W), delta)
        best = min(best, e)
    return best


def test_slackprop_pe_single_row_is_exact():
    """A 1-row grid is one chain: the bound equals the exhaustive minimum of
    the per-edge energy, and the decoded labelling attains it."""
    rng = np.random.default_rng(22)
    H, W, D = 1, 6, 3
    C = rng.normal(size=(H, W, D))
    p1x = rng.uniform(0, 1, (H, W))
    p2x = p1x + rng.uniform(0, 2, (H, W))
    zero = np.zeros((H, W))
    r = O.slackprop_pe(C, p1x, p2x, zero, zero, 2, 2)
    emin = _brute_min(C, p1x, p2x, zero, zero, 2)
    assert abs(r["lb"] - emin) < 1e-10
    assert abs(O.energy_pe(C, p1x, p2x, zero, zero, r["labels"], 2) - emin) < 1e-10


def test_slackprop_pe_grid_brackets_the_minimum():
    """3 x 3 pixels, D = 3, independent per-edge penalties (and the Potts
    form P1 = P2): LB <= min E <= E(labels) by exhaustive search."""
    rng = np.random.default_rng(23)
    H, W, D = 3, 3, 3
    C = rng.normal(size=(H, W, D))
    for potts in (False, True):
        p1x, p1y = rng.uniform(0, 1.5, (H, W)), rng.uniform(0, 1.5, (H, W))
        p2x = p1x if potts else p1x + rng.uniform(0, 2, (H, W))
        p2y = p1y if potts else p1y + rng.uniform(0, 2, (H, W))
        r = O.slackprop_pe(C, p1x, p2x, p1y, p2y, 2, 6)
        emin = _brute_min(C, p1x, p2x, p1y, p2y, 2)
        e_lab = O.energy_pe(C, p1x, p2x, p1y, p2y, r["labels"], 2)
        assert r["lb"] <= emin + 1e-10 <= e_lab + 2e-10
        seq = O.dual_sequence(r["sums"], r["kept"])
        assert all(b >= a - 1e-9 for a, b in zip(seq, seq[1:]))
