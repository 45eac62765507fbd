"""ctypes wrapper of the FP64 oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the CPU
baseline leg of bench.py as the checker.  The product path never imports it.
All functions take and return float64 numpy arrays in the layouts documented
in oracle.h (images C x H x W, cost volumes H x W x D).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)


def build():
    subprocess.check_call(["make", "-B", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        c_int, c_long, c_double = ctypes.c_int, ctypes.c_long, ctypes.c_double
        sig = {
            "or_reflect": (c_int, [c_int, c_int]),
            "or_pad_symmetric": (None, [_D, c_int, c_int, c_int, c_int, _D]),
            "or_normalize": (None, [_D, c_int, c_long, _D, _D]),
            "or_rgb2gray": (None, [_D, c_int, c_int, _D]),
            "or_conv2d": (None, [_D, c_int, c_int, c_int, _D, _D, c_int, c_int, c_int, c_int, _D]),
            "or_conv2d_bwd": (None, [_D, c_int, c_int, c_int, _D, _D, c_int, c_int, c_int, c_int,
                                     _D, _D, _D, _D]),
            "or_correlation": (None, [_D, _D, c_int, c_int, c_int, _I, c_int, c_int, _D]),
            "or_correlation_bwd": (None, [_D, _D, c_int, c_int, c_int, _I, c_int, _D, _D, _D]),
            "or_neg_softmax": (None, [_D, c_long, c_int]),
            "or_softmax_xent": (c_double, [_D, c_long, c_int, _I, _D]),
            "or_gauss": (None, [_D, c_int, c_int, c_double, _D]),
            "or_edge_weights": (None, [_D, c_int, c_int, c_double, c_double, c_double, c_double, _D, _D]),
            "or_argmin": (None, [_D, c_long, c_int, _I]),
            "or_remap": (None, [_I, c_long, c_double, c_double, _D]),
            "or_quad_fit": (None, [_D, _D, c_long, c_int, c_double, c_double]),
            "or_lr_check": (None, [_D, _D, c_int, c_int, c_double, _D]),
            "or_rho": (c_double, [c_int, c_double, c_double, c_int]),
            "or_energy": (c_double, [_D, _D, _D, c_int, c_int, c_int, _I, c_double, c_double, c_int]),
            "or_energy_masked": (c_double, [_D, _D, _D, c_int, c_int, c_int, _I, _I, c_double, c_double, c_int]),
            "or_chain_viterbi": (c_double, [_D, _D, c_int, c_int, c_double, c_double, c_int, _I]),
            "or_chain_minorant": (c_double, [_D, _D, c_int, c_int, c_double, c_double, c_int, _D, _D]),
            "or_slackprop": (c_double, [_D, _D, _D, c_int, c_int, c_int, c_double, c_double, c_int,
                                        c_int, _I, _D, _D, _D, _D]),
            "or_slackprop_pe": (c_double, [_D, _D, _D, _D, _D, c_int, c_int, c_int, c_int, c_int, _I, _D, _D, _D,
                                           _D]),
            "or_energy_pe": (c_double, [_D, _D, _D, _D, _D, c_int, c_int, c_int, _I, c_int]),
            "or_pw_relayout": (None, [_D, c_int, c_int, c_int, _D]),
            "or_ssvm_grad": (None, [_I, _I, _D, _D, c_int, c_int, c_int, c_double, c_double, c_int,
                                    _D, _D, _D, _D]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _pd(a):
    return None if a is None else a.ctypes.data_as(_D)


def _pi(a):
    return None if a is None else a.ctypes.data_as(_I)


def reflect(i, n):
    return lib().or_reflect(int(i), int(n))


def pad_symmetric(img, p):
    img = _d(img)
    C, H, W = img.shape
    out = np.empty((C, H + 2 * p, W + 2 * p))
    lib().or_pad_symmetric(_pd(img), C, H, W, p, _pd(out))
    return out


def normalize(img):
    x = _d(img).copy()
    C = x.shape[0]
    mean = np.empty(C)
    std = np.empty(C)
    lib().or_normalize(_pd(x), C, x[0].size, _pd(mean), _pd(std))
    return x, mean, std


def rgb2gray(rgb):
    rgb = _d(rgb)
    _, H, W = rgb.shape
    out = np.empty((H, W))
    lib().or_rgb2gray(_pd(rgb), H, W, _pd(out))
    return out


def conv2d(x, w, b=None, act_tanh=True):
    x, w = _d(x), _d(w)
    C, H, W = x.shape
    K, C2, KH, KW = w.shape
    assert C == C2
    b = None if b is None else _d(b)
    y = np.empty((K, H - KH + 1, W - KW + 1))
    lib().or_conv2d(_pd(x), C, H, W, _pd(w), _pd(b), K, KH, KW, int(act_tanh), _pd(y))
    return y


def conv2d_bwd(x, w, b, dy, act_tanh=True):
    x, w, dy = _d(x), _d(w), _d(dy)
    C, H, W = x.shape
    K, _, KH, KW = w.shape
    b = None if b is None else _d(b)
    dx = np.empty_like(x)
    dw = np.empty_like(w)
    db = np.empty(K)
    lib().or_conv2d_bwd(_pd(x), C, H, W, _pd(w), _pd(b), K, KH, KW, int(act_tanh), _pd(dy),
                        _pd(dx), _pd(dw), _pd(db))
    return dx, dw, db


def correlation(f0, f1, disp, rect_corr=0):
    f0, f1 = _d(f0), _d(f1)
    C, H, W = f0.shape
    disp = _i(disp)
    out = np.empty((H, W, len(disp)))
    lib().or_correlation(_pd(f0), _pd(f1), C, H, W, _pi(disp), len(disp), rect_corr, _pd(out))
    return out


def correlation_bwd(f0, f1, disp, G):
    f0, f1, G = _d(f0), _d(f1), _d(G)
    C, H, W = f0.shape
    disp = _i(disp)
    df0 = np.empty_like(f0)
    df1 = np.empty_like(f1)
    lib().or_correlation_bwd(_pd(f0), _pd(f1), C, H, W, _pi(disp), len(disp), _pd(G), _pd(df0), _pd(df1))
    return df0, df1


def neg_softmax(vol):
    v = _d(vol).copy()
    D = v.shape[-1]
    lib().or_neg_softmax(_pd(v), v.size // D, D)
    return v


def softmax_xent(S, target):
    S = _d(S)
    D = S.shape[-1]
    t = _i(target)
    dS = np.empty_like(S)
    loss = lib().or_softmax_xent(_pd(S), S.size // D, D, _pi(t), _pd(dS))
    return loss, dS


def gauss(img, sigma):
    img = _d(img)
    H, W = img.shape
    out = np.empty_like(img)
    lib().or_gauss(_pd(img), H, W, sigma, _pd(out))
    return out


def edge_weights(gray, sigma=0.7, alpha=10.0, beta=1.0, lam=1.0):
    gray = _d(gray)
    H, W = gray.shape
    wx = np.empty_like(gray)
    wy = np.empty_like(gray)
    lib().or_edge_weights(_pd(gray), H, W, sigma, alpha, beta, lam, _pd(wx), _pd(wy))
    return wx, wy


def argmin(vol):
    v = _d(vol)
    D = v.shape[-1]
    idx = np.empty(v.shape[:-1], dtype=np.int32)
    lib().or_argmin(_pd(v), v.size // D, D, _pi(idx))
    return idx


def remap(labels, disp_min, disp_step):
    lab = _i(labels)
    out = np.empty(lab.shape)
    lib().or_remap(_pi(lab), lab.size, disp_min, disp_step, _pd(out))
    return out


def quad_fit(disp, vol, disp_min, disp_step):
    """QuadDirect refinement: a refined copy of disp (H x W) from vol (H x W x D)."""
    d = _d(disp).copy()
    v = _d(vol)
    lib().or_quad_fit(_pd(d), _pd(v), d.size, v.shape[-1], disp_min, disp_step)
    return d


def lr_check(left, right, th):
    l, r = _d(left), _d(right)
    out = np.empty_like(l)
    lib().or_lr_check(_pd(l), _pd(r), l.shape[0], l.shape[1], th, _pd(out))
    return out


def rho(d, L1, L2, delta):
    return lib().or_rho(int(d), L1, L2, int(delta))


def energy(C, wx, wy, labels, L1, L2, delta):
    C, wx, wy = _d(C), _d(wx), _d(wy)
    H, W, D = C.shape
    lab = _i(labels)
    return lib().or_energy(_pd(C), _pd(wx), _pd(wy), H, W, D, _pi(lab), L1, L2, int(delta))


def energy_masked(C, wx, wy, labels, mask, L1, L2, delta):
    """Energy over the pixels whose mask label lies in [0, D) (their unaries and
    the edges between two of them)."""
    C, wx, wy = _d(C), _d(wx), _d(wy)
    H, W, D = C.shape
    lab, m = _i(labels), _i(mask)
    return lib().or_energy_masked(_pd(C), _pd(wx), _pd(wy), H, W, D, _pi(lab), _pi(m), L1, L2, int(delta))


def chain_viterbi(theta, w, L1, L2, delta):
    theta, w = _d(theta), _d(w)
    n, D = theta.shape
    lab = np.empty(n, dtype=np.int32)
    e = lib().or_chain_viterbi(_pd(theta), _pd(w), n, D, L1, L2, int(delta), _pi(lab))
    return e, lab


def chain_minorant(theta, w, L1, L2, delta):
    theta, w = _d(theta), _d(w)
    n, D = theta.shape
    g = np.empty_like(theta)
    kept = np.zeros(1)
    lam = lib().or_chain_minorant(_pd(theta), _pd(w), n, D, L1, L2, int(delta), _pd(g), _pd(kept))
    return lam, g, float(kept[0])


def slackprop(C, wx, wy, L1, L2, delta, iters, want_A=False):
    """Returns dict(labels, lb, sums, kept, gap[, A])."""
    C, wx, wy = _d(C), _d(wx), _d(wy)
    H, W, D = C.shape
    lab = np.empty((H, W), dtype=np.int32)
    sums = np.empty(2 * iters + 1)
    kept = np.empty(max(2 * iters, 1))
    gap = np.empty((H, W))
    A = np.empty_like(C) if want_A else None
    lb = lib().or_slackprop(_pd(C), _pd(wx), _pd(wy), H, W, D, L1, L2, int(delta), iters,
                            _pi(lab), _pd(sums), _pd(kept), _pd(gap), _pd(A))
    out = dict(labels=lab, lb=lb, sums=sums, kept=kept[:2 * iters], gap=gap)
    if want_A:
        out["A"] = A
    return out


def slackprop_pe(C, p1x, p2x, p1y, p2y, delta, iters, want_A=False):
    """SlackProp with per-edge penalties; returns dict(labels, lb, sums, kept, gap[, A])."""
    C = _d(C)
    p1x, p2x, p1y, p2y = _d(p1x), _d(p2x), _d(p1y), _d(p2y)
    H, W, D = C.shape
    lab = np.empty((H, W), dtype=np.int32)
    sums = np.empty(2 * iters + 1)
    kept = np.empty(max(2 * iters, 1))
    gap = np.empty((H, W))
    A = np.empty_like(C) if want_A else None
    lb = lib().or_slackprop_pe(_pd(C), _pd(p1x), _pd(p2x), _pd(p1y), _pd(p2y), H, W, D, int(delta), iters,
                               _pi(lab), _pd(sums), _pd(kept), _pd(gap), _pd(A))
    out = dict(labels=lab, lb=lb, sums=sums, kept=kept[:2 * iters], gap=gap)
    if want_A:
        out["A"] = A
    return out


def energy_pe(C, p1x, p2x, p1y, p2y, labels, delta):
    C = _d(C)
    H, W, D = C.shape
    lab = _i(labels)
    return lib().or_energy_pe(_pd(C), _pd(_d(p1x)), _pd(_d(p2x)), _pd(_d(p1y)), _pd(_d(p2y)), H, W, D, _pi(lab),
                              int(delta))


def pw_relayout(pw):
    """pw: (2 levels, H, W) -> pen (levels, H, W, 2), the reference's re-layout."""
    pw = _d(pw)
    C2, H, W = pw.shape
    pen = np.empty((C2 // 2, H, W, 2))
    lib().or_pw_relayout(_pd(pw), C2 // 2, H, W, _pd(pen))
    return pen


def pw_penalties(pw):
    """Per-edge (p1x, p2x, p1y, p2y) of a pairwise volume: P1 = pen[levels-1], P2 = pen[0]."""
    pen = pw_relayout(pw)
    L = pen.shape[0]
    return pen[L - 1, :, :, 0], pen[0, :, :, 0], pen[L - 1, :, :, 1], pen[0, :, :, 1]


def ssvm_grad(xhat, xstar, wx, wy, D, L1, L2, delta):
    xh, xs = _i(xhat), _i(xstar)
    wx, wy = _d(wx), _d(wy)
    H, W = xh.shape
    dC = np.empty((H, W, D))
    dL = np.empty(2)
    dwx = np.empty((H, W))
    dwy = np.empty((H, W))
    lib().or_ssvm_grad(_pi(xh), _pi(xs), _pd(wx), _pd(wy), H, W, D, L1, L2, int(delta),
                       _pd(dC), _pd(dL), _pd(dwx), _pd(dwy))
    return dC, dL, dwx, dwy


def dual_sequence(sums, kept):
    """Dual bound before the first half-step and after each half-step:
    [S_1, K_1 + S_2, K_2 + S_3, ...] (monotone non-decreasing)."""
    return [sums[0]] + [kept[j - 1] + sums[j] for j in range(1, len(sums))]
